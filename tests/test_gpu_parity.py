"""GPU parity: librte.so (HIP, gfx950) against the FP64 oracle and the
reference CPU package's golden outputs, on identical seeded inputs.

Tolerances (max |gpu - ref| / max |ref|):
  FP64 path            1e-10 (operators) / 1e-9 (full solves)
  mixed FP32/FP64 path 1e-5 (operators, solves)
  plain FP32 path      1e-4 (solves)
"""
import numpy as np
import pytest

from doi_amd import rte
from oracle import oracle as O
from tests.helpers import bounds_for, hetero_maps, rel_max

pytestmark = pytest.mark.gpu

PREC_TOL = {rte.PREC_FP64: 1e-10, rte.PREC_MIXED: 1e-5, rte.PREC_FP32: 2e-5}


def _field(K, N, seed):
    return np.random.default_rng(seed).standard_normal((K, N))


@pytest.mark.parametrize("shape,L", [((4, 4, 4), 0), ((4, 4, 4), 3), ((5, 3, 4), 2), ((4, 5, 6), 5),
                                     ((6, 6, 6), 1), ((4, 4, 3), 6), ((3, 4, 4), 7)])
@pytest.mark.parametrize("prec", [rte.PREC_FP64, rte.PREC_MIXED, rte.PREC_FP32])
@pytest.mark.parametrize("path", [rte.PATH_CACHED, rte.PATH_MARCH])
def test_propagate_heterogeneous(ctx, shape, L, prec, path):
    print("propagate", shape, L, prec, path, flush=True)
    bounds = bounds_for(shape, d=(0.1, 0.12, 0.09))
    ma, ms = hetero_maps(shape, seed=sum(shape) + L)
    K, N = (L + 1) ** 2, int(np.prod(shape))
    W = _field(K, N, seed=7 + L)
    P = O.Problem(shape, bounds, ma, ms, 0.8, 1.3)
    ref = P.propagate(W, L)
    got = ctx.propagate(shape, bounds, ma, ms, 0.8, W, L, precision=prec, path_mode=path, n=1.3)
    assert rel_max(got, ref) < PREC_TOL[prec]


@pytest.mark.parametrize("thr", [0.15, 0.25])
def test_propagate_threshold(ctx, thr):
    shape, L = (5, 5, 5), 2
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=3)
    W = _field(9, 125, seed=11)
    ref = O.Problem(shape, bounds, ma, ms, 0.5).propagate(W, L, prop_thresh=thr)
    for path in (rte.PATH_CACHED, rte.PATH_MARCH):
        got = ctx.propagate(shape, bounds, ma, ms, 0.5, W, L, prop_thresh=thr, precision=rte.PREC_FP64,
                            path_mode=path)
        assert rel_max(got, ref) < 1e-10


def test_propagate_uniform_self_block(ctx):
    """Uniform medium: the oracle uses the reference's direct 200x200 self-block
    quadrature, the GPU its moment series."""
    shape, L = (4, 4, 4), 5
    bounds = bounds_for(shape)
    W = _field(36, 64, seed=5)
    ref = O.Problem(shape, bounds, 0.01, 1.0, 0.8).propagate(W, L)
    got = ctx.propagate(shape, bounds, 0.01, 1.0, 0.8, W, L, precision=rte.PREC_FP64)
    assert rel_max(got, ref) < 1e-10


def test_source_and_ballistic(ctx):
    shape, L = (6, 5, 4), 4
    bounds = bounds_for(shape, d=(0.1, 0.1, 0.15))
    ma, ms = hetero_maps(shape, seed=9)
    beam = (0.05, 0.0, 0.3, 1.7)
    s0, ball, cols = O.Problem(shape, bounds, ma, ms, 0.7).source(beam, L)
    g0, gball, gcols = ctx.source(shape, bounds, ma, ms, 0.7, beam, L)
    assert gcols == cols > 0
    assert rel_max(g0, s0) < 1e-13
    assert rel_max(gball, ball) < 1e-13


def test_scatter_and_exitance(ctx):
    shape, L = (5, 4, 3), 3
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=4)
    W = _field(16, 60, seed=2)
    P = O.Problem(shape, bounds, ma, ms, 0.6, 1.4)
    assert rel_max(ctx.scatter(shape, bounds, ma, ms, 0.6, W, L, n=1.4), P.scatter(W, L)) < 1e-14
    for q in (rte.QUAD_REF, rte.QUAD_EXACT):
        got = ctx.exitance(shape, bounds, W, L, q, c_m=2.9979e10 / 1.4)
        assert rel_max(got, P.exitance(W, L, q)) < 1e-13


def test_path_integrals(ctx):
    shape = (7, 6, 5)
    bounds = bounds_for(shape, d=(0.1, 0.08, 0.12))
    ma, ms = hetero_maps(shape, seed=1, sigma=0.5)
    rng = np.random.default_rng(0)
    N = int(np.prod(shape))
    d = rng.integers(0, N, 400)
    s = rng.integers(0, N, 400)
    P = O.Problem(shape, bounds, ma, ms, 0.0)
    ref = np.array([P.tau(int(a), int(b)) for a, b in zip(d, s)])
    got = ctx.path_integrals(shape, bounds, ma, ms, d, s)
    assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-30)) < 2e-6


@pytest.mark.parametrize("prec,tol", [(rte.PREC_FP64, 1e-9), (rte.PREC_MIXED, 1e-5), (rte.PREC_FP32, 1e-4)])
def test_golden_reference_cases(ctx, golden_ref, prec, tol):
    """Every uniform case the reference binary produced: same term count, and
    the transmitted and reflected images within the precision's tolerance."""
    for c in golden_ref:
        b = c["beam"]
        p = rte.make_params(c["L"], precision=prec, stop_rule=rte.STOP_LAMBDA,
                            prop_thresh=c["prop_thresh"])
        print("golden", c["name"], prec, flush=True)
        m = ctx.medium(tuple(c["shape"]), tuple(c["bounds"]), c["mua"], c["mus"], c["g"], n=c["n"])
        imgs, refls, sts = ctx.solve_images([m], [(b["pos_x"], b["pos_y"], b["diam"], b["mag"])], p)
        m.close()
        assert sts[0]["terms"] == c["terms"], c["name"]
        assert rel_max(imgs[0], np.array(c["trans_image"])) < tol, c["name"]
        assert rel_max(refls[0], np.array(c["refl_image"])) < tol, c["name"]


@pytest.mark.parametrize("rule", [rte.STOP_LAMBDA, rte.STOP_DELTA])
@pytest.mark.parametrize("prec,tol", [(rte.PREC_FP64, 1e-9), (rte.PREC_MIXED, 1e-5)])
@pytest.mark.parametrize("path", [rte.PATH_CACHED, rte.PATH_MARCH])
def test_heterogeneous_solve(ctx, rule, prec, tol, path):
    shape, L = (8, 8, 8), 3
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=21, sigma=0.3)
    beam = (0.05, -0.05, 0.1, 1.0)
    ref = O.Problem(shape, bounds, ma, ms, 0.8).solve(beam, L, stop_rule=rule, tol=1e-6, max_terms=30)
    p = rte.make_params(L, precision=prec, stop_rule=rule, tol=1e-6, max_terms=30, path_mode=path)
    img, st = ctx.solve(shape, bounds, ma, ms, 0.8, beam, p)
    assert st["terms"] == ref["terms"]
    assert rel_max(img, ref["exitance"]) < tol


def test_batch_equals_alone(ctx):
    """A solve gives the same bits alone and inside a batch of other solves."""
    shape, L = (8, 8, 8), 3
    bounds = bounds_for(shape)
    p = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    mm = [ctx.medium(shape, bounds, *hetero_maps(shape, seed=s), 0.8) for s in (1, 2, 3)]
    beams = [(0.05, 0.05, 0.1, 1.0), (-0.15, 0.05, 0.1, 1.0), (0.0, 0.0, 0.25, 1.0)]
    alone = [ctx.solve_media([m], [b], p)[0][0] for m, b in zip(mm, beams)]
    batch, _ = ctx.solve_media(mm + [mm[0]], beams + [beams[1]], p)
    for i in range(3):
        assert np.array_equal(batch[i], alone[i])
    other = ctx.solve_media([mm[0]], [beams[1]], p)[0][0]
    assert np.array_equal(batch[3], other)
    for m in mm:
        m.close()


DEFAULT_GROUP = 5  # rte_ctx_set_group_size's default


@pytest.mark.parametrize("L", [3, 4, 5, 6, 7])
def test_beams_sharing_a_medium_equal_alone(ctx, L):
    """Seven beams on one medium, propagated as groups of 5 and 2 (group size 5,
    the default), 4 and 3 (group size 4), 3, 3 and 1 (group size 3) or 2, 2, 2
    and 1 (group size 2) right-hand sides, give the same bits as each beam solved
    alone (L 3 and 5: the group of three reads W rows in halves; L 4: in whole
    rows; groups of three to five prefetch their rows; L 6 and 7: the library
    caps the groups at 3 and 2)."""
    shape = (12, 12, 12)
    bounds = bounds_for(shape)
    m = ctx.medium(shape, bounds, *hetero_maps(shape, seed=31), 0.9)
    xs = (np.arange(12) + 0.5) * 0.1 - 0.6
    beams = [(xs[i], xs[(3 * i + 2) % 12], 0.1, 1.0) for i in range(7)]
    p = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=20)
    alone = [ctx.solve_media([m], [b], p) for b in beams]
    try:
        for gsize in (5, 4, 3, 2):
            ctx.set_group_size(gsize)
            batch, st = ctx.solve_media([m] * 7, beams, p)
            for i in range(7):
                assert alone[i][1][0]["terms"] == st[i]["terms"]
                assert np.array_equal(alone[i][0][0], batch[i]), (gsize, i)
    finally:
        ctx.set_group_size(DEFAULT_GROUP)
    m.close()


def test_solves_on_different_media_equal_alone(ctx):
    """Five media with one beam each in one batch, at the default group size
    (pairs across media) and group size 1: every result has the same bits as
    the solve alone."""
    shape, L = (12, 12, 12), 5
    bounds = bounds_for(shape)
    mm = [ctx.medium(shape, bounds, *hetero_maps(shape, seed=40 + k), 0.85) for k in range(5)]
    xs = (np.arange(12) + 0.5) * 0.1 - 0.6
    beams = [(xs[2 * k + 1], xs[11 - 2 * k], 0.1, 1.0) for k in range(5)]
    p = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=25)
    alone = [ctx.solve_media([m], [b], p)[0][0] for m, b in zip(mm, beams)]
    batch, _ = ctx.solve_media(mm, beams, p)
    try:
        ctx.set_group_size(1)
        unpaired, _ = ctx.solve_media(mm, beams, p)
    finally:
        ctx.set_group_size(DEFAULT_GROUP)
    for k in range(5):
        assert np.array_equal(batch[k], alone[k]), k
        assert np.array_equal(unpaired[k], alone[k]), k
    for m in mm:
        m.close()


@pytest.mark.parametrize("L", [6, 7])
def test_high_order_solve(ctx, L):
    """The two highest supported orders end to end against the oracle (FP64 and mixed)."""
    shape = (6, 6, 5)
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=60 + L)
    beam = (0.05, -0.05, 0.1, 1.0)
    ref = O.Problem(shape, bounds, ma, ms, 0.85).solve(beam, L, stop_rule=O.STOP_DELTA, tol=1e-6, max_terms=20)
    for prec, tol in ((rte.PREC_FP64, 1e-9), (rte.PREC_MIXED, 1e-5)):
        p = rte.make_params(L, precision=prec, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=20)
        img, st = ctx.solve(shape, bounds, ma, ms, 0.85, beam, p)
        assert st["terms"] == ref["terms"]
        assert rel_max(img, ref["exitance"]) < tol
