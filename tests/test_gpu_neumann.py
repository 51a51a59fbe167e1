"""GPU parity of rte_neumann, the reference's Neumann(src, out, 1) on a given
first term, against the oracle's orc_neumann (pinned in
tests/test_oracle_neumann.py); the beam's own source through rte_neumann
against rte_solve_field; and the FIM scan (doi_amd.fim) against the same scan
composed from the oracle.

Tolerances (max |gpu - ref| / max |ref|): FP64 path 1e-9, mixed FP32/FP64 path
1e-5; the Fisher information, a sum of squares over a ratio, 1e-8 and 1e-4.
"""
import numpy as np
import pytest

from doi_amd import fim, rte, snr
from oracle import oracle as O
from tests.helpers import bounds_for, hetero_maps, rel_max

pytestmark = pytest.mark.gpu

TOL = {rte.PREC_FP64: 1e-9, rte.PREC_MIXED: 1e-5}


@pytest.mark.parametrize("prec", [rte.PREC_FP64, rte.PREC_MIXED])
@pytest.mark.parametrize("rule,path", [(rte.STOP_LAMBDA, rte.PATH_CACHED), (rte.STOP_DELTA, rte.PATH_CACHED),
                                       (rte.STOP_LAMBDA, rte.PATH_MARCH)])
def test_neumann_against_oracle(ctx, prec, rule, path):
    shape, L = (6, 5, 4), 3
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=21, mua=0.02, mus=1.3)
    K, N = (L + 1) ** 2, int(np.prod(shape))
    rng = np.random.default_rng(5)
    src = rng.random((K, N)) * np.exp(-np.arange(K))[:, None] - 0.1 * rng.random((K, N))
    print("neumann", prec, rule, path, flush=True)
    w, img, refl, terms = O.Problem(shape, bounds, ma, ms, 0.75, n=1.3).neumann(
        src, L, stop_rule=rule, tol=1e-8, max_terms=40)
    m = ctx.medium(shape, bounds, ma, ms, 0.75, n=1.3)
    p = rte.make_params(L, precision=prec, stop_rule=rule, tol=1e-8, max_terms=40, path_mode=path)
    f, gi, gr, st = ctx.neumann(m, src, p)
    m.close()
    assert st["terms"] == terms
    assert rel_max(f, w) < TOL[prec]
    assert rel_max(gi, img) < TOL[prec]
    assert rel_max(gr, refl) < TOL[prec]


def test_beam_source_through_neumann_is_the_beam_solve(ctx):
    shape, L = (7, 6, 5), 4
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=4, mua=0.01, mus=0.9)
    beam = (0.03, 0.01, 0.25, 1.5)
    s0, ball, _ = ctx.source(shape, bounds, ma, ms, 0.85, beam, L, n=1.4)
    m = ctx.medium(shape, bounds, ma, ms, 0.85, n=1.4)
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_LAMBDA)
    field, img, refl, st = ctx.solve_field(m, beam, p)
    f, gi, gr, st2 = ctx.neumann(m, s0, p)
    m.close()
    assert st2["terms"] == st["terms"]
    assert rel_max(f, field) < 1e-12
    assert rel_max(gi + np.reshape(ball, gi.shape), img) < 1e-12
    assert rel_max(gr, refl) < 1e-12


def test_neumann_errors(ctx):
    shape, L = (4, 4, 3), 2
    ma, ms = hetero_maps(shape, seed=1)
    m = ctx.medium(shape, bounds_for(shape), ma, ms, 0.8)
    src = np.ones(((L + 1) ** 2, m.N))
    with pytest.raises(rte.RteError):
        ctx.neumann(m, src, rte.make_params(L, precision=rte.PREC_FP32))
    src[3, 5] = np.nan
    with pytest.raises(rte.RteError):
        ctx.neumann(m, src, rte.make_params(L, precision=rte.PREC_FP64))
    m.close()


@pytest.mark.parametrize("prec,tol", [(rte.PREC_FP64, 1e-12), (rte.PREC_MIXED, 1e-6)])
@pytest.mark.parametrize("path", [rte.PATH_CACHED, rte.PATH_MARCH])
def test_neumann_reference_layout(ctx, prec, tol, path):
    """rte_neumann_ref (the reference's Neumann(src, out, flag) on its complex,
    voxel-major field) against rte_neumann on the same first term in the real
    basis (flag 1: same term count, the field mapped by to_reference_layout) and
    against rte_propagate (flag 0: out <- A out); fields that are not a real
    radiance, and flag 1 with out not preset to src, are rejected."""
    shape, L = (6, 5, 4), 3
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=23, mua=0.02, mus=1.2)
    K, N = (L + 1) ** 2, int(np.prod(shape))
    rng = np.random.default_rng(9)
    src = rng.random((K, N)) * np.exp(-np.arange(K))[:, None] - 0.05 * rng.random((K, N))
    csrc = rte.to_reference_layout(src, L)
    m = ctx.medium(shape, bounds, ma, ms, 0.8, n=1.37)
    p = rte.make_params(L, precision=prec, stop_rule=rte.STOP_LAMBDA, path_mode=path)
    print("neumann_ref flag 1", prec, path, flush=True)
    f, _, _, st = ctx.neumann(m, src, p)
    out, st2 = ctx.neumann_ref(m, csrc, p, flag=1)
    assert st2["terms"] == st["terms"]
    ref = rte.to_reference_layout(f, L)
    assert np.max(np.abs(out - ref)) < tol * np.max(np.abs(ref))
    assert rel_max(rte.from_reference_layout(out, L), f) < tol
    print("neumann_ref flag 0", prec, path, flush=True)
    a0, _ = ctx.neumann_ref(m, None, p, flag=0, out=csrc)
    aw = ctx.propagate(shape, bounds, ma, ms, 0.8, src, L, precision=prec, path_mode=path, n=1.37)
    ref0 = rte.to_reference_layout(aw, L)
    assert np.max(np.abs(a0 - ref0)) < tol * np.max(np.abs(ref0))
    bad = csrc.copy()
    bad[3, 1] += 0.5j          # Im c_10 != 0: not a real radiance
    with pytest.raises(rte.RteError) as ei:
        ctx.neumann_ref(m, bad, p, flag=1)
    assert ei.value.code == rte.E_ARG
    with pytest.raises(rte.RteError) as ei:
        ctx.neumann_ref(m, csrc, p, flag=1, out=2.0 * csrc)
    assert ei.value.code == rte.E_ARG
    m.close()


@pytest.mark.parametrize("prec,tol", [(rte.PREC_FP64, 1e-8), (rte.PREC_MIXED, 1e-4)])
def test_fim_scan_against_oracle(ctx, prec, tol):
    n, a, L = 8, 0.4, 3
    shape, bounds = (n, n, n), (-a, a, -a, a, 0.0, 2 * a)
    N = n ** 3
    ma, ms = np.full(N, 0.01), np.full(N, 1.0)     # every tissue as rte_fim.cu:226-229 sets it
    ma[N // 2 + 3] = 0.05                           # and one voxel that is not
    beam, nr, c = (0.0, 0.0, 0.3, 2.0), 1.37, 2.9979e10
    depths, radii = (1, 4), fim.reference_radii(shape, bounds)[:3]
    p = rte.make_params(L, precision=prec, stop_rule=rte.STOP_LAMBDA)
    print("fim scan", prec, flush=True)
    res = fim.fim_scan(ctx, shape, bounds, ma, ms, 0.85, beam, p, n=nr, c=c, depths=depths, radii=radii)
    P = O.Problem(shape, bounds, ma, ms, 0.85, n=nr, c=c)
    w = P.solve(beam, L, field=True)["w"]
    g0 = snr.fim_top_image(w, shape, bounds, L, c=c, n=nr, mag=beam[3])
    assert rel_max(res["g"], g0) < TOL[prec]
    for i, z in enumerate(depths):
        for j, r in enumerate(radii):
            inside = fim.region_mask(shape, bounds, (n // 2, n // 2, z), r)
            dw, _, _, _ = P.neumann(np.where(inside[None, :], -c * w, 0.0), L)
            dg = snr.fim_top_image(dw, shape, bounds, L, c=c, n=nr, mag=beam[3])
            F = fim.fisher_top(g0, dg)
            assert abs(res["fisher"][i, j] - F) <= tol * abs(F), (z, r)


def test_fim_command_line(ctx, golden_ref, tmp_path, capsys):
    """python -m doi_amd.fim file.ini prints the reference's line with the
    value fim_scan gives for the file's problem (z = 1, the first radius)."""
    from doi_amd import refio
    from tests.test_refio import ini_for
    case = next(c for c in golden_ref if c["shape"][0] == c["shape"][1] == c["shape"][2])
    p = tmp_path / "fim.ini"
    p.write_text(ini_for(case))
    assert fim.main([str(p), "--precision", "fp64"]) == 0
    line = capsys.readouterr().out.strip().splitlines()[-1]
    prob = refio.load(str(p))
    params = rte.make_params(prob.L, precision=rte.PREC_FP64, stop_rule=rte.STOP_LAMBDA, max_terms=1000,
                             prop_thresh=prob.prop_thresh)
    res = fim.fim_scan(ctx, prob.shape, prob.bounds, prob.mu_a, prob.mu_s, prob.g, prob.beam, params, n=prob.n,
                       c=prob.c)
    r = fim.reference_radii(prob.shape, prob.bounds)[0]
    assert line == "%e is fisher abs for %d as z index and %f as radius" % (res["fisher"][0, 0], 1, r)
