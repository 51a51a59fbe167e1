"""The oracle's Neumann(src, out, 1) on a given first term (orc_neumann), pinned
to the beam solve that the reference binary's golden images pin (the beam's s0
as the first term gives the same series, field and image less the ballistic
term), to the reference's forward-mode derivative (a derivative source through
orc_neumann gives orc_ref_derivative's image), and by linearity in the source;
and the FIM scan's region and Fisher sum on the reference's own formulas
(CPU only)."""
import numpy as np

from doi_amd import fim
from oracle import oracle as O
from tests.helpers import bounds_for, hetero_maps, rel_max

SHAPE, L = (5, 4, 4), 3


def _problem(seed=3, n=1.4):
    ma, ms = hetero_maps(SHAPE, seed, mua=0.03, mus=1.1)
    return O.Problem(SHAPE, bounds_for(SHAPE), ma, ms, 0.8, n=n), (0.02, -0.01, 0.2, 1.3)


def test_neumann_of_the_beam_source_is_the_beam_solve():
    P, beam = _problem()
    for rule in (O.STOP_LAMBDA, O.STOP_DELTA):
        r = P.solve(beam, L, stop_rule=rule, tol=1e-8, field=True)
        s0, ball, _ = P.source(beam, L)
        w, img, _, terms = P.neumann(s0, L, stop_rule=rule, tol=1e-8)
        assert terms == r["terms"]
        assert np.array_equal(w, r["w"])
        nx, ny, _ = SHAPE
        assert rel_max(img + ball.reshape(ny, nx), r["exitance"]) < 1e-14


def test_neumann_carries_the_reference_derivative_and_is_linear():
    P, beam = _problem(seed=8)
    region = np.zeros(P.N, dtype=np.int32)
    region[[17, 18, 37]] = 1
    phi = P.solve(beam, L, field=True)["w"]
    src = np.where(region[None, :] != 0, -P.optics.c * phi, 0.0)   # -1.0*C, compute_grad_rte.cu:292
    _, img, _, terms = P.neumann(src, L)
    ref, ref_terms = P.ref_derivative(beam, L, region, 0)
    assert terms == ref_terms
    assert rel_max(img, ref) < 1e-14
    # a fixed number of terms: the field is linear in the first term
    rng = np.random.default_rng(2)
    s1, s2 = rng.random(src.shape), rng.random(src.shape) - 0.5
    kw = dict(stop_rule=O.STOP_DELTA, tol=0.0, max_terms=6)
    w1, w2, w12 = (P.neumann(s, L, **kw)[0] for s in (s1, s2, 2.0 * s1 - 3.0 * s2))
    assert rel_max(w12, 2.0 * w1 - 3.0 * w2) < 1e-12


def test_fim_region_and_fisher_sum_follow_the_reference_formulas():
    # the reference's coordinates (rte_fim.cu:173-179) on a cube [-a, a]^2 x [0, 2a]
    n, a = 8, 0.4
    shape, bounds = (n, n, n), (-a, a, -a, a, 0.0, 2 * a)
    fov = 2 * a
    x = np.arange(n) * fov / n - fov / 2 + fov / (2 * n)
    z = np.arange(n) * fov / n + fov / (2 * n)
    radii = fim.reference_radii(shape, bounds)
    assert np.allclose(radii, [(i + 1) * fov / (2 * n) for i in range(5)])
    for k, r in [(1, radii[0]), (3, radii[2]), (2, radii[4])]:
        keep = np.zeros((n, n, n), dtype=bool)
        for iz in range(n):
            for iy in range(n):
                for ix in range(n):
                    d2 = (x[ix] - x[n // 2]) ** 2 + (x[iy] - x[n // 2]) ** 2 + (z[iz] - z[k]) ** 2
                    keep[iz, iy, ix] = not d2 > r * r
        assert np.array_equal(fim.region_mask(shape, bounds, (n // 2, n // 2, k), r), keep.ravel())
    # the smallest radius holds the centre voxel alone
    assert fim.region_mask(shape, bounds, (4, 4, 1), radii[0]).sum() == 1
    g = np.array([[2.0, 0.0], [4.0, -1.0]])
    dg = np.array([[1.0, 5.0], [2.0, 3.0]])
    assert fim.fisher_top(g, dg) == 1.0 / 2.0 + 4.0 / 4.0 + 9.0 / -1.0
