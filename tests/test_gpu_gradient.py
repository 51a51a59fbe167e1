"""Reverse-mode derivatives of the exitance (rte_exitance_vjp / rte_gradient)
against central finite differences of the forward solve with the Neumann
term count held fixed -- on the GPU in FP64 and on the FP64 oracle -- plus
checkpointing and precision invariants.

Tolerance: |adjoint - FD| <= 2e-6 * max|FD| + 1e-5 |FD| (FP64, step 1e-4 relative)."""
import numpy as np
import pytest

from doi_amd import rte
from oracle import oracle as O
from tests.helpers import bounds_for, hetero_maps

pytestmark = pytest.mark.gpu

SHAPE, L, G = (6, 6, 5), 3, 0.8
BEAM = (0.05, -0.05, 0.1, 1.0)


def _setup(seed=11):
    bounds = bounds_for(SHAPE)
    ma, ms = hetero_maps(SHAPE, seed=seed, sigma=0.3)
    cot = np.random.default_rng(seed).standard_normal(SHAPE[0] * SHAPE[1])
    return bounds, ma, ms, cot


def _fd(fun, ma, ms, which, v, h_rel=1e-4):
    a, s = ma.copy(), ms.copy()
    base = (a if which == "a" else s)[v]
    h = h_rel * base
    (a if which == "a" else s)[v] = base + h
    fp = fun(a, s)
    (a if which == "a" else s)[v] = base - h
    fm = fun(a, s)
    return (fp - fm) / (2 * h)


def _close(adj, fd, scale):
    return abs(adj - fd) <= 2e-6 * scale + 1e-5 * abs(fd)


def test_vjp_matches_finite_differences_gpu_and_oracle(ctx):
    bounds, ma, ms, cot = _setup()
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-8, max_terms=40)
    med = ctx.medium(SHAPE, bounds, ma, ms, G)
    img, ga, gs, st = ctx.exitance_vjp(med, BEAM, p, cot)
    n = st["terms"]
    med.close()
    pf = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=0.0, max_terms=n)

    def gpu_fun(a, s):
        e, stt = ctx.solve(SHAPE, bounds, a, s, G, BEAM, pf)
        assert stt["terms"] == n
        return float(np.dot(cot, e.ravel()))

    P = O.Problem(SHAPE, bounds, ma, ms, G)

    def orc_fun(a, s):
        r = O.Problem(SHAPE, bounds, a, s, G).solve(BEAM, L, stop_rule=O.STOP_DELTA, tol=0.0, max_terms=n)
        return float(np.dot(cot, r["exitance"].ravel()))

    N = int(np.prod(SHAPE))
    rng = np.random.default_rng(3)
    # a beam column voxel, a bottom voxel, and random ones
    vox = [2 + 6 * 2 + 36 * 1, 3 + 6 * 3 + 36 * 4] + list(rng.choice(N, 8, replace=False))
    scale = max(np.max(np.abs(ga)), np.max(np.abs(gs)))
    for v in vox:
        for which, adj in (("a", ga[v]), ("s", gs[v])):
            fd_g = _fd(gpu_fun, ma, ms, which, v)
            assert _close(adj, fd_g, scale), (v, which, adj, fd_g)
    for v in vox[:4]:
        for which, adj in (("a", ga[v]), ("s", gs[v])):
            fd_o = _fd(orc_fun, ma, ms, which, v)
            assert _close(adj, fd_o, scale), (v, which, adj, fd_o)
    # the image of the vjp call is the forward solve's
    ref = P.solve(BEAM, L, stop_rule=O.STOP_DELTA, tol=1e-8, max_terms=40)
    assert np.max(np.abs(img - ref["exitance"])) < 1e-9 * np.max(np.abs(ref["exitance"]))


def test_checkpoint_stride_is_exact(ctx):
    """Recomputing the terms between checkpoints reproduces them exactly; the
    back-projection adds with atomics, so the gradients agree to rounding."""
    bounds, ma, ms, cot = _setup(seed=5)
    med = ctx.medium(SHAPE, bounds, ma, ms, G)
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_LAMBDA)
    _, ga1, gs1, st1 = ctx.exitance_vjp(med, BEAM, p, cot, checkpoint_stride=1)
    for stride in (2, 3, 7):
        _, ga, gs, st = ctx.exitance_vjp(med, BEAM, p, cot, checkpoint_stride=stride)
        assert st["terms"] == st1["terms"]
        sc = max(np.max(np.abs(ga1)), np.max(np.abs(gs1)))
        assert np.max(np.abs(ga - ga1)) < 1e-12 * sc and np.max(np.abs(gs - gs1)) < 1e-12 * sc
    med.close()


def test_loss_gradient_and_mixed_precision(ctx):
    bounds, ma, ms, _ = _setup(seed=8)
    ma2, ms2 = hetero_maps(SHAPE, seed=99, sigma=0.3)
    med = ctx.medium(SHAPE, bounds, ma, ms, G)
    p64 = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-7, max_terms=30)
    target, _ = ctx.solve(SHAPE, bounds, ma2, ms2, G, BEAM, p64)
    loss, img, ga, gs, st = ctx.gradient(med, BEAM, p64, target)
    assert abs(loss - np.sum((img - target) ** 2)) < 1e-12 * max(loss, 1e-30)
    _, ga2, gs2, _ = ctx.exitance_vjp(med, BEAM, p64, 2 * (img - target))
    sc0 = max(np.max(np.abs(ga)), np.max(np.abs(gs)))
    assert np.max(np.abs(ga - ga2)) < 1e-12 * sc0 and np.max(np.abs(gs - gs2)) < 1e-12 * sc0
    pm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-7, max_terms=30)
    lm, _, gam, gsm, stm = ctx.gradient(med, BEAM, pm, target)
    assert stm["terms"] == st["terms"]
    sc = max(np.max(np.abs(ga)), np.max(np.abs(gs)))
    assert np.max(np.abs(gam - ga)) < 1e-4 * sc and np.max(np.abs(gsm - gs)) < 1e-4 * sc
    med.close()


def test_fisher_information(ctx):
    """J against central differences of block parameters, F = J^T diag(1/E) J
    (FP32 MFMA) against an FP64 host Gram of the returned J, Cholesky factor."""
    shape = (6, 6, 5)
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=17, sigma=0.3)
    # 3 x 3 x 2 blocks of 2 x 2 x (3|2) voxels -> 18 parameters (not a multiple of 16)
    i, j, k = np.meshgrid(np.arange(6), np.arange(6), np.arange(5), indexing="ij")
    part = ((i // 2) + 3 * (j // 2) + 9 * (k // 3)).transpose(2, 1, 0).ravel()
    nparam = 18
    med = ctx.medium(shape, bounds, ma, ms, G)
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-8, max_terms=40)
    F, Cf, J, img, st = ctx.fisher(med, BEAM, p, part, nparam, wrt=0)
    n = st["terms"]
    med.close()
    E = img.ravel()
    assert np.all(E > 0)
    # J columns by central differences (n fixed)
    pf = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=0.0, max_terms=n)
    for a in (0, 4, 13, 17):
        h = 1e-4 * np.mean(ma)
        up, dn = ma.copy(), ma.copy()
        up[part == a] += h
        dn[part == a] -= h
        ep, _ = ctx.solve(shape, bounds, up, ms, G, BEAM, pf)
        em, _ = ctx.solve(shape, bounds, dn, ms, G, BEAM, pf)
        fd = (ep - em).ravel() / (2 * h)
        assert np.max(np.abs(J[:, a] - fd)) < 2e-6 * np.max(np.abs(J)) + 1e-5 * np.max(np.abs(fd))
    # Gram on MFMA against FP64 on the host
    F64 = J.T @ (J / E[:, None])
    assert np.max(np.abs(F - F64)) < 2e-5 * np.max(np.abs(F64))
    assert np.array_equal(F, F.T)
    assert np.allclose(np.tril(Cf), Cf)
    assert np.max(np.abs(Cf @ Cf.T - F)) < 1e-12 * np.max(np.abs(F))


@pytest.mark.parametrize("wrt", [0, 1])
def test_exact_jacobian_rows_are_block_sums_of_the_vjp(ctx, wrt):
    """The exact Fisher path back-projects onto the parameter blocks directly, two
    detectors per backward pass; each row of its J must equal the per-voxel
    vector-Jacobian product of that detector summed over the blocks (FP64, a
    partition with voxels outside every block; 25 detectors, so the last pass has one)."""
    shape = (5, 5, 4)
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=23, sigma=0.3)
    i, j, k = np.meshgrid(np.arange(5), np.arange(5), np.arange(4), indexing="ij")
    part = ((i // 2) + 3 * (j // 2) + 9 * (k // 2)).transpose(2, 1, 0).ravel()
    part[(np.arange(part.size) % 7) == 3] = -1
    nparam = int(part.max()) + 1
    beam = (0.0, 0.0, 0.1, 1.0)  # on the centre column of the odd grid
    med = ctx.medium(shape, bounds, ma, ms, G)
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-8, max_terms=40)
    _, _, J, img, _ = ctx.fisher(med, beam, p, part, nparam, wrt=wrt, jacobian=0)
    E = img.ravel()
    for pix in (0, 7, 12, 24):
        if E[pix] == 0:  # a detector the Jacobian skips
            continue
        cot = np.zeros(E.size)
        cot[pix] = 1.0
        _, ga, gs, _ = ctx.exitance_vjp(med, beam, p, cot)
        g = ga if wrt == 0 else gs
        row = np.array([g[part == a].sum() for a in range(nparam)])
        assert np.max(np.abs(J[pix] - row)) < 1e-10 * np.max(np.abs(row)), pix
    med.close()


@pytest.mark.parametrize("nref", [1.0, 1.37])
def test_reference_jacobian_matches_oracle(ctx, nref):
    """rte_ref_jacobian (the reference's own forward-mode derivative) against the
    oracle's restatement of it, with every series held at the same term count:
    FP64 within 1e-9 and mixed precision within 1e-4 of max |J|; Fisher built on
    it.  n = 1.37 exercises the reference's -1.0*C factor, which differs from
    -c_m by n.  The gap to the exact derivative is reported, not bounded: the
    reference approximates dA/dmu by -A chi A."""
    shape = (6, 6, 5)
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=23, sigma=0.3)
    i, j, k = np.meshgrid(np.arange(6), np.arange(6), np.arange(5), indexing="ij")
    part = ((i // 3) + 2 * (j // 3) + 4 * (k // 3)).transpose(2, 1, 0).ravel()  # 8 blocks
    nparam, n = 8, 12
    med = ctx.medium(shape, bounds, ma, ms, G, n=nref)
    P = O.Problem(shape, bounds, ma, ms, G, n=nref)
    for wrt in (0, 1):
        p64 = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=0.0, max_terms=n)
        J, img, st = ctx.ref_jacobian(med, BEAM, p64, part, nparam, wrt)
        assert st["terms"] == n
        sc = np.max(np.abs(J))
        for a in range(nparam):
            ref, terms = P.ref_derivative(BEAM, L, (part == a).astype(np.int32), wrt, stop_rule=O.STOP_DELTA,
                                          tol=0.0, max_terms=n)
            assert terms == n
            assert np.max(np.abs(J[:, a] - ref.ravel())) < 1e-9 * sc, (wrt, a)
        pm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=0.0, max_terms=n)
        Jm, _, _ = ctx.ref_jacobian(med, BEAM, pm, part, nparam, wrt)
        assert np.max(np.abs(Jm - J)) < 1e-4 * sc
        if nref != 1.0:
            continue
        # exact derivative of the same blocks, for the record
        _, _, Jx, _, _ = ctx.fisher(med, BEAM, p64, part, nparam, wrt=wrt, jacobian=0)
        gap = np.linalg.norm(J - Jx) / np.linalg.norm(Jx)
        print("wrt %d: |J_ref - J_exact| / |J_exact| = %.3f" % (wrt, gap))
        assert np.isfinite(gap)
    if nref != 1.0:
        med.close()
        return
    F, Cf, Jf, img, _ = ctx.fisher(med, BEAM, p64, part, nparam, wrt=0, jacobian=1)
    J0, _, _ = ctx.ref_jacobian(med, BEAM, p64, part, nparam, 0)
    assert np.max(np.abs(Jf - J0)) <= 1e-12 * np.max(np.abs(J0))
    E = img.ravel()
    F64 = Jf.T @ (Jf / E[:, None])
    assert np.max(np.abs(F - F64)) < 2e-5 * np.max(np.abs(F64))
    med.close()


def test_vjp_beyond_the_lds_accumulator_28cube(ctx):
    """At 28^3 (21952 voxels: the FP64 gradient no longer fits one workgroup's
    160 KB of LDS) the back-projection adds into device memory with global
    atomics (k_backproject_sym<..., false>); the adjoint still matches central
    differences of the forward solve at fixed term count, FP64, L = 1."""
    shape, L = (28, 28, 28), 1
    assert shape[0] * shape[1] * shape[2] * 8 > 160 * 1024
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=28, sigma=0.3)
    cot = np.random.default_rng(28).standard_normal(shape[0] * shape[1])
    n = 4
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=0.0, max_terms=n)
    beam = (0.05, -0.05, 0.1, 1.0)
    med = ctx.medium(shape, bounds, ma, ms, G)
    print("28^3 vjp", flush=True)
    _, ga, gs, st = ctx.exitance_vjp(med, beam, p, cot)
    med.close()
    assert st["terms"] == n

    def gpu_fun(a, s):
        e, stt = ctx.solve(shape, bounds, a, s, G, beam, p)
        return float(np.dot(cot, e.ravel()))

    nx = shape[0]
    # below the beam, in the middle, at the bottom and off to the side
    vox = [14 + nx * 13 + nx * nx * 3, 14 + nx * 14 + nx * nx * 14, 10 + nx * 15 + nx * nx * 27,
           3 + nx * 22 + nx * nx * 9]
    scale = max(np.max(np.abs(ga)), np.max(np.abs(gs)))
    for v in vox:
        for which, adj in (("a", ga[v]), ("s", gs[v])):
            print("28^3 fd", v, which, flush=True)
            fd_g = _fd(gpu_fun, ma, ms, which, v)
            assert _close(adj, fd_g, scale), (v, which, adj, fd_g)
