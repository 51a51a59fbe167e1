"""GPU parity at the workloads' full sizes: the 16^3 L=5 lognormal medium against the
committed FP64 oracle fixture, and size-independent properties at 32^3
(mixed vs FP64 error bar, mirror symmetry, linearity in the beam power,
batch independence, the frequency-domain solve's symmetries) and of rte_neumann
(the beam's own first term at 24^3, linearity in the first term at 32^3)."""
import json
import os

import numpy as np
import pytest

from doi_amd import media, rte
from tests.helpers import rel_max

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lognormal_16cube_fixture():
    with open(os.path.join(HERE, "golden", "oracle_lognormal_16cube_L5.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("prec,tol", [(rte.PREC_FP64, 1e-9), (rte.PREC_MIXED, 1e-5), (rte.PREC_FP32, 1e-4)])
def test_lognormal_16cube_against_oracle_fixture(ctx, lognormal_16cube_fixture, prec, tol):
    fx = lognormal_16cube_fixture
    c = media.lognormal_16cube(seed=fx["seed"])
    for case in fx["cases"]:
        print("16^3 L=5", case["name"], prec, flush=True)
        p = rte.make_params(fx["L"], precision=prec, stop_rule=case["stop_rule"], tol=case["tol"],
                            max_terms=case["max_terms"])
        img, st = ctx.solve(c["shape"], c["bounds"], c["mu_a"], c["mu_s"], c["g"], c["beam"], p)
        assert st["terms"] == case["terms"]
        assert rel_max(img, np.array(case["exitance"])) < tol


def test_32cube_mixed_within_1e4_of_fp64(ctx):
    n, L = 32, 5
    ma, ms = media.lognormal_media_with_inclusions(1, seed=5, n=n)[0]
    m = ctx.medium((n, n, n), media.cube_bounds(n), ma, ms, 0.9)
    beams = media.top_face_beams(seed=5, n=n)[:2]
    out = {}
    for prec in (rte.PREC_FP64, rte.PREC_MIXED, rte.PREC_FP32):
        print("32^3", prec, flush=True)
        p = rte.make_params(L, precision=prec, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=12)
        out[prec] = ctx.solve_media([m, m], beams, p)
    for prec in (rte.PREC_MIXED, rte.PREC_FP32):
        for b in range(2):
            assert out[prec][1][b]["terms"] == out[rte.PREC_FP64][1][b]["terms"]
            assert rel_max(out[prec][0][b], out[rte.PREC_FP64][0][b]) < 1e-4
    m.close()


def test_32cube_five_beams_default_groups(ctx):
    """The benchmark's own propagation instance: five beams on one 32^3 L=5
    medium of the benchmark's media, mixed precision, default group size (one
    group of five right-hand sides) and group size 3 (groups of three and two),
    tolerance 1e-6 with a cap of 30 terms.  Same term counts as the all-FP64
    solves, within 1e-4 of them, and the same bits as each beam solved alone."""
    n, L = 32, 5
    ma, ms = media.lognormal_media_with_inclusions(1, seed=5, n=n)[0]
    m = ctx.medium((n, n, n), media.cube_bounds(n), ma, ms, 0.9)
    beams = media.top_face_beams(seed=5, n=n)
    assert len(beams) == 5
    pm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    p64 = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    print("32^3 five beams, one group of 5", flush=True)
    batch, st = ctx.solve_media([m] * 5, beams, pm)
    print("32^3 five beams, groups of 3 + 2", flush=True)
    try:
        ctx.set_group_size(3)
        batch32, st32 = ctx.solve_media([m] * 5, beams, pm)
    finally:
        ctx.set_group_size(5)
    print("32^3 five beams FP64", flush=True)
    ref, st64 = ctx.solve_media([m] * 5, beams, p64)
    print("32^3 five beams alone", flush=True)
    for b in range(5):
        alone, sta = ctx.solve_media([m], [beams[b]], pm)
        assert sta[0]["terms"] == st[b]["terms"] == st32[b]["terms"] == st64[b]["terms"], b
        assert np.array_equal(alone[0], batch[b]), b
        assert np.array_equal(alone[0], batch32[b]), b
        assert rel_max(batch[b], ref[b]) < 1e-4, b
    m.close()


def test_32cube_symmetry_linearity_batch(ctx):
    n, L = 32, 3
    # x- and y-mirror-symmetric heterogeneous medium, centred beam (4 columns)
    rng = np.random.default_rng(4)
    q = 0.7 + 0.6 * rng.random((n, n // 2, n // 2))
    full = np.concatenate([q, q[:, :, ::-1]], axis=2)
    full = np.concatenate([full, full[:, ::-1, :]], axis=1).ravel()
    bounds = media.cube_bounds(n)
    m = ctx.medium((n, n, n), bounds, 0.01 * full, full, 0.85)
    other = ctx.medium((n, n, n), bounds, *media.heterogeneous(n, 9, sigma=0.3), 0.85)
    p = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=8)
    imgs, st = ctx.solve_media([m, m, other], [(0, 0, 0.2, 1.0), (0, 0, 0.2, 2.5), (0.35, -0.25, 0.1, 1.0)], p)
    e = imgs[0]
    assert rel_max(e[:, ::-1], e) < 1e-5
    assert rel_max(e[::-1, :], e) < 1e-5
    assert rel_max(imgs[1], 2.5 * e) < 1e-6
    alone, _ = ctx.solve_media([other], [(0.35, -0.25, 0.1, 1.0)], p)
    assert np.array_equal(alone[0], imgs[2])
    m.close()
    other.close()


def test_32cube_frequency_domain_properties(ctx):
    """The frequency-domain solve at 32^3, L = 5 (100 MHz): the mixed path within
    1e-4 of the FP64 one, -omega the complex conjugate, the modulus linear in the
    beam power and omega -> 0 the continuous-wave image."""
    n, L, om = 32, 5, 2.0 * np.pi * 1.0e8
    ma, ms = media.lognormal_media_with_inclusions(1, seed=5, n=n)[0]
    m = ctx.medium((n, n, n), media.cube_bounds(n), ma, ms, 0.9)
    b = media.top_face_beams(seed=5, n=n)[0]
    out = {}
    for prec in (rte.PREC_FP64, rte.PREC_MIXED):
        print("32^3 fd", prec, flush=True)
        p = rte.make_params(L, precision=prec, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=12)
        out[prec] = ctx.solve_fd([m], [b], [om], p)
    e64, e32 = out[rte.PREC_FP64][0][0], out[rte.PREC_MIXED][0][0]
    assert out[rte.PREC_MIXED][2][0]["terms"] == out[rte.PREC_FP64][2][0]["terms"]
    assert np.max(np.abs(e32 - e64)) < 1e-4 * np.max(np.abs(e64))
    p = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=12)
    neg, _, _ = ctx.solve_fd([m], [b], [-om], p)
    assert np.max(np.abs(neg[0] - np.conj(e32))) <= 1e-14 * np.max(np.abs(e32))
    twice, _, _ = ctx.solve_fd([m], [(b[0], b[1], b[2], 2.0)], [om], p)
    assert np.max(np.abs(twice[0] - 2.0 * e32)) < 1e-6 * np.max(np.abs(e32))
    cw, _ = ctx.solve_media([m], [b], p)
    slow, _, _ = ctx.solve_fd([m], [b], [1.0], p)       # 1 rad/s: phase ~1e-10
    assert np.max(np.abs(slow[0] - cw[0])) < 1e-6 * np.max(np.abs(cw[0]))
    m.close()


def test_neumann_at_full_size(ctx):
    """rte_neumann at the workloads' sizes: at 24^3, L = 5 (FP64) the beam's own
    first term gives rte_solve_field's field and image; at 32^3, L = 5 (mixed,
    six terms) the field is linear in the first term."""
    n, L = 24, 5
    ma, ms = media.lognormal_media_with_inclusions(1, seed=7, n=n)[0]
    bounds = media.cube_bounds(n)
    beam = media.top_face_beams(seed=7, n=n)[0]
    print("neumann 24^3 beam source", flush=True)
    s0, ball, _ = ctx.source((n, n, n), bounds, ma, ms, 0.9, beam, L)
    m = ctx.medium((n, n, n), bounds, ma, ms, 0.9)
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=12)
    field, img, _, st = ctx.solve_field(m, beam, p)
    f, gi, _, st2 = ctx.neumann(m, s0, p)
    m.close()
    assert st2["terms"] == st["terms"]
    assert rel_max(f, field) < 1e-12
    assert rel_max(gi + np.reshape(ball, gi.shape), img) < 1e-12
    n = 32
    ma, ms = media.lognormal_media_with_inclusions(1, seed=8, n=n)[0]
    m = ctx.medium((n, n, n), media.cube_bounds(n), ma, ms, 0.9)
    rng = np.random.default_rng(3)
    K = (L + 1) ** 2
    s1 = rng.random((K, m.N)) * np.exp(-np.arange(K))[:, None]
    s2 = rng.random((K, m.N)) * np.exp(-np.arange(K))[:, None]
    p = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=0.0, max_terms=6)
    print("neumann 32^3 linearity", flush=True)
    f1, f2, f12 = (ctx.neumann(m, s, p)[0] for s in (s1, s2, s1 + 2.0 * s2))
    m.close()
    assert rel_max(f12, f1 + 2.0 * f2) < 1e-5
