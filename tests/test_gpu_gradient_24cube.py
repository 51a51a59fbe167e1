"""Differentiation and Fisher information at full size: a heterogeneous 24^3
grid, L=5, maps as in the 16^3 parity case (lognormal sigma 0.2, one radius-3
sphere with mu_a x2), loss = squared error against the exitance of a second
seeded map.

* Adjoint gradients (mixed precision: FP32 propagation, FP64 sums) with respect
  to per-voxel mu_a and mu_s against central finite differences (step 1e-4
  relative) of FP64 forward solves at the same term count, on 32 random voxels:
  |adjoint - FD| <= 1e-4 max|FD| + 1e-3 |FD|.
* Fisher information for 64 parameters (mu_a on 4x4x4 blocks of 6^3 voxels) and
  576 bottom detectors, J by the reference's method: F = J^T diag(1/E) J on FP32
  MFMA is symmetric, matches an FP64 Gram of the returned J, and its FP64
  Cholesky factor reconstructs it."""
import time

import numpy as np
import pytest

from doi_amd import media, rte

pytestmark = pytest.mark.gpu

n, L, G = 24, 5, 0.8
SHAPE = (n, n, n)
BOUNDS = media.cube_bounds(n)
BEAM = (0.05, 0.05, 0.2, 1.0)


def _maps(seed):
    return media.heterogeneous(n, seed, sigma=0.2, n_incl=1, radius=(3, 3), factor=(2.0, 2.0))


def test_adjoint_gradient_24cube_against_finite_differences(ctx):
    ma, ms = _maps(3)
    ma2, ms2 = _maps(4)
    p64 = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    target, _ = ctx.solve(SHAPE, BOUNDS, ma2, ms2, G, BEAM, p64)
    med = ctx.medium(SHAPE, BOUNDS, ma, ms, G)
    pm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    t0 = time.perf_counter()
    loss, img, ga, gs, st = ctx.gradient(med, BEAM, pm, target)
    t_adj = time.perf_counter() - t0
    med.close()
    nt = st["terms"]
    pf = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=0.0, max_terms=nt)

    def loss_of(a, s):
        e, stt = ctx.solve(SHAPE, BOUNDS, a, s, G, BEAM, pf)
        assert stt["terms"] == nt
        return float(np.sum((e - target) ** 2))

    rng = np.random.default_rng(33)
    vox = rng.choice(n ** 3, 32, replace=False)
    fd_a, fd_s = [], []
    t0 = time.perf_counter()
    for v in vox:
        for which, out in (("a", fd_a), ("s", fd_s)):
            a, s = ma.copy(), ms.copy()
            base = (a if which == "a" else s)[v]
            h = 1e-4 * base
            (a if which == "a" else s)[v] = base + h
            lp = loss_of(a, s)
            (a if which == "a" else s)[v] = base - h
            lm = loss_of(a, s)
            out.append((lp - lm) / (2 * h))
    t_fd = time.perf_counter() - t0
    fd_a, fd_s = np.array(fd_a), np.array(fd_s)
    print("24^3 gradient: %d terms, adjoint %.2f s for all %d voxels x 2; FD %.2f s for 32 voxels x 2"
          % (nt, t_adj, n ** 3, t_fd))
    for adj, fd in ((ga[vox], fd_a), (gs[vox], fd_s)):
        sc = np.max(np.abs(fd))
        assert np.all(np.abs(adj - fd) <= 1e-4 * sc + 1e-3 * np.abs(fd)), np.max(np.abs(adj - fd)) / sc


def test_fisher_24cube_64_parameters_576_detectors(ctx):
    ma, ms = _maps(3)
    med = ctx.medium(SHAPE, BOUNDS, ma, ms, G)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    part = ((i // 6) + 4 * (j // 6) + 16 * (k // 6)).transpose(2, 1, 0).ravel()
    pm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    # the reference's Jacobian (one derivative series per parameter, batched); the
    # exact one costs one adjoint series per detector and is tested at 6x6x5
    for jac in (1,):
        t0 = time.perf_counter()
        F, Cf, J, img, st = ctx.fisher(med, BEAM, pm, part, 64, wrt=0, jacobian=jac)
        dt = time.perf_counter() - t0
        print("24^3 Fisher (%s Jacobian): %.2f s, %d terms" % ("reference" if jac else "exact", dt, st["terms"]))
        E = img.ravel()
        assert J.shape == (576, 64) and np.all(E > 0)
        F64 = J.T @ (J / E[:, None])
        assert np.array_equal(F, F.T)
        assert np.max(np.abs(F - F64)) < 2e-5 * np.max(np.abs(F64))
        assert np.max(np.abs(Cf @ Cf.T - F)) < 1e-10 * np.max(np.abs(F))
    med.close()


def test_exact_fisher_24cube_columns_against_finite_differences(ctx):
    """The exact Jacobian at full size (reverse mode, one adjoint series per
    detector, mixed precision): four of its 64 columns (mu_a on 6^3 blocks, the
    beam's block, a deep one, a bottom one and a far corner) against central
    differences of FP64 forward solves at the same term count:
    |J - FD| <= 2e-3 max|FD| per column.  (The reference's forward-mode
    Jacobian misses this by tens of percent.)  F is formed from it on MFMA."""
    ma, ms = _maps(3)
    med = ctx.medium(SHAPE, BOUNDS, ma, ms, G)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    part = ((i // 6) + 4 * (j // 6) + 16 * (k // 6)).transpose(2, 1, 0).ravel()
    pm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    print("24^3 exact Fisher", flush=True)
    t0 = time.perf_counter()
    F, Cf, J, img, st = ctx.fisher(med, BEAM, pm, part, 64, wrt=0, jacobian=0)
    dt = time.perf_counter() - t0
    med.close()
    nt = st["terms"]
    print("24^3 Fisher (exact Jacobian, 576 adjoint series): %.2f s, %d terms" % (dt, nt), flush=True)
    E = img.ravel()
    assert J.shape == (576, 64) and np.all(E > 0)
    F64 = J.T @ (J / E[:, None])
    assert np.max(np.abs(F - F64)) < 2e-5 * np.max(np.abs(F64))
    # F^-1 in FP64 from the Cholesky factor (the Cramer-Rao bound's matrix)
    Fi = rte.fisher_inverse(Cf)
    assert np.max(np.abs(F @ Fi - np.eye(64))) < 1e-6
    assert np.all(np.diag(Fi) > 0)
    print("Cramer-Rao bound on mu_a per block: %.3g .. %.3g cm^-1" % tuple(np.sqrt(np.diag(Fi))[[0, -1]]))
    pf = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=0.0, max_terms=nt)
    for a in (1 + 4 * 1, 1 + 4 * 1 + 16 * 2, 2 + 4 * 2 + 16 * 3, 3 + 4 * 0 + 16 * 1):
        blk = part == a
        h = 1e-3 * float(np.mean(ma[blk]))
        up, dn = ma.copy(), ma.copy()
        up[blk] += h
        dn[blk] -= h
        print("24^3 FD column", a, flush=True)
        ep, s1 = ctx.solve(SHAPE, BOUNDS, up, ms, G, BEAM, pf)
        em, s2 = ctx.solve(SHAPE, BOUNDS, dn, ms, G, BEAM, pf)
        assert s1["terms"] == s2["terms"] == nt
        fd = (ep - em).ravel() / (2 * h)
        err = np.max(np.abs(J[:, a] - fd)) / np.max(np.abs(fd))
        print("column %d: max |J - FD| / max |FD| = %.2e" % (a, err), flush=True)
        assert err < 2e-3, (a, err)
