"""Pins the FP64 oracle (oracle/rte_oracle.c) before it is trusted as the
checker: against the reference CPU package's own outputs (built from its
sources by oracle/Makefile.ref, fixtures in tests/golden/ref_cpu_uniform.json)
and, for what those leave open (heterogeneous media, the real basis, the
moment series), against closed forms, library routines, invariants and brute
force.  CPU only.
"""
import json
import os

import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O
from tests.helpers import bounds_for, hetero_maps, rel_max

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
# the reference binary's outputs (uniform media)
# ---------------------------------------------------------------------------
def test_reference_golden_cases(golden_ref):
    """Transmitted and reflected images and term counts of every case the
    reference binary produced."""
    for c in golden_ref:
        P = O.Problem(tuple(c["shape"]), tuple(c["bounds"]), c["mua"], c["mus"], c["g"], c["n"])
        b = c["beam"]
        r = P.solve((b["pos_x"], b["pos_y"], b["diam"], b["mag"]), c["L"], prop_thresh=c["prop_thresh"],
                    reflectance=True)
        assert r["terms"] == c["terms"], c["name"]
        assert rel_max(r["exitance"], np.array(c["trans_image"])) < 1e-12, c["name"]
        assert rel_max(r["reflectance"], np.array(c["refl_image"])) < 1e-12, c["name"]


def test_reflectance_quadrature():
    """int_{-1}^0 P_l(u) u du = -1/2, 1/3, -1/8, 0, 1/48, 0 (exact mode), and the
    reference's 100-point left Riemann sum over u = p/100 - 1 in closed form for
    l = 0, 1: sum u / 100 = -0.505, sum u^2 / 100 = 0.33835."""
    R = O.quad_weights_refl(5, O.QUAD_EXACT)
    assert np.allclose(R, [-0.5, 1 / 3, -0.125, 0.0, 1 / 48, 0.0], atol=1e-14)
    Rr = O.quad_weights_refl(5, O.QUAD_REF)
    assert abs(Rr[0] + 0.505) < 1e-15 and abs(Rr[1] - 0.33835) < 1e-15
    # reflectance of a top-layer field is the exitance functional on the mirrored grid
    shape = (3, 2, 4)
    bounds = bounds_for(shape)
    P = O.Problem(shape, bounds, 0.01, 1.0, 0.5)
    L, K, N = 3, 16, 24
    w = np.random.default_rng(2).standard_normal((K, N))
    flip = w.reshape(K, 4, 6)[:, ::-1, :].reshape(K, N) * np.array(
        [(-1.0) ** int(np.sqrt(a + 0.5)) for a in range(K)])[:, None]   # z -> -z: R_l0 picks up (-1)^l
    for q in (O.QUAD_REF, O.QUAD_EXACT):
        refl = P.reflectance(w, L, q)
        ex = P.exitance(flip, L, q)
        if q == O.QUAD_EXACT:
            assert rel_max(refl, ex) < 1e-13


# ---------------------------------------------------------------------------
# spherical harmonics
# ---------------------------------------------------------------------------
ANGLES = [(0.0, 0.0), (0.3, 0.7), (1.2, -2.5), (np.pi / 2, np.pi / 4), (2.9, 3.1), (np.pi, 1.0)]


def test_ylm_closed_forms():
    for th, ph in ANGLES:
        y = O.ylm_complex(2, th, ph)
        c, s, e = np.cos(th), np.sin(th), np.exp(1j * ph)
        exact = {0: 0.5 / np.sqrt(np.pi),
                 2: np.sqrt(3 / (4 * np.pi)) * c,
                 3: -np.sqrt(3 / (8 * np.pi)) * s * e,
                 6: np.sqrt(5 / (16 * np.pi)) * (3 * c * c - 1),
                 7: -np.sqrt(15 / (8 * np.pi)) * s * c * e,
                 8: np.sqrt(15 / (32 * np.pi)) * s * s * e * e}
        for a, v in exact.items():
            assert abs(y[a] - v) < 1e-14


def test_ylm_against_scipy_and_symmetry():
    L = 7
    for th, ph in ANGLES:
        y = O.ylm_complex(L, th, ph)
        r = O.ylm_real(L, th, ph)
        for l in range(L + 1):
            for m in range(-l, l + 1):
                a = l * l + l + m
                assert abs(y[a] - sp.sph_harm_y(l, m, th, ph)) < 1e-12
                if m > 0:
                    assert abs(y[l * l + l - m] - (-1) ** m * np.conj(y[a])) < 1e-14
                    assert abs(r[a] - np.sqrt(2) * y[a].real) < 1e-14
                    assert abs(r[l * l + l - m] - np.sqrt(2) * y[a].imag) < 1e-14
                elif m == 0:
                    assert abs(r[a] - y[a].real) < 1e-14


def test_real_basis_orthonormal():
    L = 6
    xg, wg = np.polynomial.legendre.leggauss(16)
    nphi = 32
    K = (L + 1) ** 2
    G = np.zeros((K, K))
    for x, w in zip(xg, wg):
        for k in range(nphi):
            R = O.ylm_real(L, np.arccos(x), 2 * np.pi * k / nphi)
            G += w * (2 * np.pi / nphi) * np.outer(R, R)
    assert np.max(np.abs(G - np.eye(K))) < 1e-12


# ---------------------------------------------------------------------------
# exitance quadrature
# ---------------------------------------------------------------------------
def test_quadrature_constants():
    Q = O.quad_weights(5, O.QUAD_REF)
    # the reference's 100-point left Riemann sum of int_0^1 P_l(u) u du
    assert np.allclose(Q, [0.495, 0.32835, 0.1200375, -0.0049415, -0.0257446, -0.0048672], atol=5e-7)
    u = np.arange(100) / 100.0
    for l in range(6):
        assert abs(Q[l] - np.sum(sp.eval_legendre(l, u) * u) / 100.0) < 1e-15
    E = O.quad_weights(7, O.QUAD_EXACT)
    assert np.allclose(E[:6], [0.5, 1 / 3, 1 / 8, 0.0, -1 / 48, 0.0], atol=1e-15)
    assert abs(E[6] - 1 / 128) < 1e-15   # int_0^1 x P_6 = 1/128


# ---------------------------------------------------------------------------
# Siddon path integral
# ---------------------------------------------------------------------------
def _brute_tau(shape, bounds, mut, d, s, nsamp=200000):
    nx, ny, nz = shape
    dx = (bounds[1] - bounds[0]) / nx
    dy = (bounds[3] - bounds[2]) / ny
    dz = (bounds[5] - bounds[4]) / nz

    def centre(v):
        i, j, k = v % nx, (v // nx) % ny, v // (nx * ny)
        return np.array([bounds[0] + (i + 0.5) * dx, bounds[2] + (j + 0.5) * dy, bounds[4] + (k + 0.5) * dz])

    a, b = centre(s), centre(d)
    t = (np.arange(nsamp) + 0.5) / nsamp
    p = a[None, :] + t[:, None] * (b - a)[None, :]
    i = np.clip(np.floor((p[:, 0] - bounds[0]) / dx).astype(int), 0, nx - 1)
    j = np.clip(np.floor((p[:, 1] - bounds[2]) / dy).astype(int), 0, ny - 1)
    k = np.clip(np.floor((p[:, 2] - bounds[4]) / dz).astype(int), 0, nz - 1)
    return np.linalg.norm(b - a) * np.mean(mut[i + nx * (j + ny * k)])


def test_tau_uniform_symmetric_and_brute_force():
    shape = (5, 4, 6)
    bounds = bounds_for(shape, d=(0.1, 0.13, 0.08))
    rng = np.random.default_rng(3)
    N = int(np.prod(shape))
    ma, ms = hetero_maps(shape, seed=4, sigma=0.6)
    P = O.Problem(shape, bounds, ma, ms, 0.0)
    U = O.Problem(shape, bounds, 0.2, 1.3, 0.0)
    mut = ma + ms
    for _ in range(40):
        d, s = (int(x) for x in rng.integers(0, N, 2))
        if d == s:
            continue
        tu = U.tau(d, s)
        r = tu / 1.5
        assert abs(tu - 1.5 * r) < 1e-14
        # uniform: mu_t * r
        nx, ny = shape[0], shape[1]
        dv = np.array([d % nx - s % nx, (d // nx) % ny - (s // nx) % ny, d // (nx * ny) - s // (nx * ny)])
        rr = np.linalg.norm(dv * np.array([0.1, 0.13, 0.08]))
        assert abs(tu - 1.5 * rr) < 1e-13
        # reciprocity and brute force
        t1, t2 = P.tau(d, s), P.tau(s, d)
        assert abs(t1 - t2) < 1e-12 * t1
        assert abs(t1 - _brute_tau(shape, bounds, mut, d, s)) < 2e-3 * t1


def test_tau_layered_exact():
    """mu_t varying in z only: every crossing order gives sum_k mu_k * (z-length in layer k) * r/|dz|."""
    shape = (4, 4, 5)
    bounds = bounds_for(shape)
    layer = np.array([0.5, 1.0, 2.0, 0.7, 1.3])
    mut = np.repeat(layer, 16)
    P = O.Problem(shape, bounds, mut, np.zeros(80), 0.0)
    d = 1 + 4 * 2 + 16 * 4   # (1,2,4)
    s = 3 + 4 * 0 + 16 * 0   # (3,0,0)
    r = np.linalg.norm([0.2, 0.2, 0.4])
    # z from 0.05 to 0.45: half of layer 0, layers 1-3, half of layer 4
    expect = r / 0.4 * (0.05 * layer[0] + 0.1 * (layer[1] + layer[2] + layer[3]) + 0.05 * layer[4])
    assert abs(P.tau(d, s) - expect) < 1e-13


# ---------------------------------------------------------------------------
# self-voxel block
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mu", [0.01, 1.01, 6.0])
def test_self_block_moment_series_equals_direct(mu):
    L = 4
    d = (0.1, 0.1, 0.1)
    A = O.self_block(L, *d, mu, 1.0, 0)
    B = O.self_block(L, *d, mu, 1.0, 1)
    assert rel_max(B, A) < 1e-12
    assert np.max(np.abs(A - A.T)) < 1e-12 * np.max(np.abs(A))


def test_self_block_limits_and_selection_rules():
    L = 4
    S = O.self_block(L, 0.1, 0.1, 0.1, 1e-9, 1.0, 1)
    beta = O.self_beta(0.1, 0.1, 0.1)
    # mu -> 0: (1 - e^{-mu b})/mu -> b, so S_00 -> dth dph/dV sum_w b sin(theta) / (4 pi)
    th = np.repeat(np.arange(200) * np.pi / 200, 200)
    s00 = (np.pi / 200) * (2 * np.pi / 200) / 1e-3 * np.sum(beta * np.sin(th)) / (4 * np.pi)
    assert abs(S[0, 0] - s00) < 1e-6 * s00
    # the mean chord from the centre of a cube of edge h: sum b dOmega / 4pi ~ 0.5 * 1.1... h/2 < b < h sqrt3/2
    assert 0.05 < s00 * 1e-3 < 0.0866
    # parity: l + l' odd couples nothing for a centred cube (to float quirks of beta)
    ls = np.array([int(np.floor(np.sqrt(a))) for a in range(25)])
    odd = (ls[:, None] + ls[None, :]) % 2 == 1
    assert np.max(np.abs(S[odd])) < 1e-6 * np.max(np.abs(S))
    # the cube's fourfold axis: m - m' in the cos/sin real basis only couples like with like
    assert np.max(np.abs(S[0, 1:4])) < 1e-6 * abs(S[0, 0])


def test_beta_is_the_reference_float_geometry():
    beta = O.self_beta(0.1, 0.1, 0.1)
    assert beta.shape == (40000,)
    assert np.all(beta >= 0.05 - 1e-9) and np.all(beta <= 0.0866026)
    # straight down the z axis the face is 0.05 away
    assert abs(beta[0] - 0.05) < 1e-12


# ---------------------------------------------------------------------------
# source, scatter, operator structure
# ---------------------------------------------------------------------------
def test_source_uniform_closed_form():
    shape, L = (6, 6, 5), 3
    bounds = bounds_for(shape, d=(0.1, 0.1, 0.12))
    mua, mus, g = 0.05, 1.2, 0.7
    P = O.Problem(shape, bounds, mua, mus, g)
    s0, ball, cols = P.source((0.0, 0.0, 0.2, 2.0), L)
    assert cols == 4
    mt, dz, dV = mua + mus, 0.12, 0.1 * 0.1 * 0.12
    z = (np.arange(5)) * dz
    v = 2 + 6 * 2 + 36 * np.arange(5)     # column (2,2)
    for l in range(L + 1):
        expect = 0.5 * g ** l * (mus / mt) * np.sqrt((2 * l + 1) / (4 * np.pi)) * \
            (np.exp(-mt * z) - np.exp(-mt * (z + dz))) / dV
        assert rel_max(s0[l * l + l, v], expect) < 1e-13
    assert abs(ball[2 + 6 * 2] - 0.5 * np.exp(-mt * 0.6)) < 1e-15
    assert np.count_nonzero(ball) == 4


def test_reciprocity_of_propagation():
    """A^T = P A P off the diagonal voxel blocks (P = (-1)^l): tau and r are
    symmetric and R(-u) = (-1)^l R(u)."""
    shape, L = (3, 2, 2), 2
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=8, sigma=0.5)
    P = O.Problem(shape, bounds, ma, ms, 0.8)
    K, N = 9, 12
    A = np.zeros((K * N, K * N))
    for col in range(K * N):
        e = np.zeros(K * N)
        e[col] = 1.0
        A[:, col] = P.propagate(e.reshape(K, N), L).ravel()
    ls = np.repeat([0, 1, 1, 1, 2, 2, 2, 2, 2], N)
    Pm = np.diag((-1.0) ** ls)
    vox = np.tile(np.arange(N), K)
    off = vox[:, None] != vox[None, :]
    D = (A.T - Pm @ A @ Pm)[off]
    assert np.max(np.abs(D)) < 1e-12 * np.max(np.abs(A))


def test_exitance_independent_of_c_and_n():
    shape, L = (5, 5, 4), 2
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=5)
    a = O.Problem(shape, bounds, ma, ms, 0.8, 1.0, 2.9979e10).solve((0, 0, 0.2, 1), L)
    b = O.Problem(shape, bounds, ma, ms, 0.8, 1.4, 1.0).solve((0, 0, 0.2, 1), L)
    assert a["terms"] == b["terms"]
    assert rel_max(b["exitance"], a["exitance"]) < 1e-13


def test_pure_absorber_is_ballistic_only():
    shape = (4, 4, 4)
    bounds = bounds_for(shape)
    P = O.Problem(shape, bounds, 0.7, 0.0, 0.5)
    r = P.solve((0, 0, 0.2, 1.0), 2)
    e = r["exitance"]
    assert r["terms"] == 2
    assert abs(e.sum() - np.exp(-0.7 * 0.4)) < 1e-14
    assert np.count_nonzero(e) == 4


def test_lambda_loop_equals_sum_of_term_exitances():
    """The reference computes w = A (sum_q W_q) after the loop; by linearity of A
    that equals the sum of the exitances of A W_q, which is how the GPU driver
    accumulates.  Force the total-exitance rule to run the same terms."""
    shape, L = (6, 6, 6), 3
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=6)
    P = O.Problem(shape, bounds, ma, ms, 0.8)
    lam = P.solve((0.05, 0.05, 0.1, 1.0), L)
    dl = P.solve((0.05, 0.05, 0.1, 1.0), L, stop_rule=O.STOP_DELTA, tol=0.0, max_terms=lam["terms"])
    assert dl["terms"] == lam["terms"]
    assert rel_max(dl["exitance"], lam["exitance"]) < 1e-13


def test_mirror_symmetry_of_centred_beam():
    """A medium symmetric under x -> -x and y -> -y with a centred beam gives a
    mirror-symmetric exitance image."""
    n = 6
    shape = (n, n, 5)
    bounds = bounds_for(shape)
    rng = np.random.default_rng(2)
    half = 0.5 + rng.random((5, n // 2, n // 2))
    full = np.concatenate([half, half[:, :, ::-1]], axis=2)
    full = np.concatenate([full, full[:, ::-1, :]], axis=1)
    P = O.Problem(shape, bounds, 0.02 * full.ravel(), full.ravel(), 0.8)
    e = P.solve((0.0, 0.0, 0.2, 1.0), 3)["exitance"]
    assert rel_max(e[:, ::-1], e) < 1e-12
    assert rel_max(e[::-1, :], e) < 1e-12


def test_lognormal_16cube_fixture_reproduces():
    """The committed 16^3 oracle fixture still matches the oracle (guards the
    fixture the GPU tests use) -- run on a subsampled stop rule to stay fast."""
    with open(os.path.join(HERE, "golden", "oracle_lognormal_16cube_L5.json")) as f:
        fx = json.load(f)
    from doi_amd import media
    c = media.lognormal_16cube(seed=fx["seed"])
    assert list(c["shape"]) == fx["shape"]
    lam = [x for x in fx["cases"] if x["name"] == "lambda_ref"][0]
    assert lam["terms"] == 10
    assert len(lam["exitance"]) == 16 and len(lam["exitance"][0]) == 16


def _series_exitance(P, s0, L, n, delta=None, eps=0.0):
    """Exitance without the ballistic term of the forward series with the
    scatter D replaced by D + eps Delta, from the oracle's separately pinned
    source, propagation, scatter and exitance:  W_0 = s0, W_{q+1} = (D + eps
    Delta) A W_q (n terms after W_0), E = e(A sum_q W_q)."""
    W = s0.copy()
    out = s0.copy()
    for _ in range(n):
        AW = P.propagate(W, L)
        W = P.scatter(AW, L)
        if delta is not None:
            W = W + eps * delta(AW)
        out += W
    return P.exitance(P.propagate(out, L), L).reshape(P.shape[1], P.shape[0])


def _reference_factor(P, L, which):
    """Per-coefficient factor of the reference's derivative source on its field
    out2 (reference units): -1.0*C for mu_a (compute_grad_rte.cu:292) and
    g^l out2 - C out2 for mu_s (PropScatmu1, GPU package/src/dist.cu:675-686, plus
    scale_dist(out3, -1.0*C), compute_grad_rte.cu:350-353), C the vacuum speed
    (FisherInfo_IO/src/rte.h:15), the optics' c here."""
    K = (L + 1) ** 2
    ls = np.floor(np.sqrt(np.arange(K) + 0.5)).astype(int)
    C = P.optics.c
    return -C * np.ones(K) if which == 0 else P.optics.g ** ls * C - C


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("nref", [1.0, 1.37])
def test_reference_derivative_is_the_scatter_perturbation(which, nref):
    """The reference's forward-mode derivative dE = e(A sum_j (DA)^j Delta phi),
    phi = A sum_q (DA)^q s0, is the derivative in eps of the forward series
    with D replaced by D + eps Delta, Delta = the reference's factor on the
    region (A and s0 held fixed).  Central differences in eps of that series,
    built from the pinned operators, must give orc_ref_derivative's image: a
    wrong field (phi vs sum W), a missing A, a dropped region or a wrong
    factor all fail.  The series converge to 1e-15 within the term count."""
    shape, L, g = (4, 4, 5), 2, 0.75
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=33, mua=0.2, mus=1.5, sigma=0.3)
    P = O.Problem(shape, bounds, ma, ms, g, n=nref)
    beam = (0.0, 0.05, 0.1, 1.0)
    N = int(np.prod(shape))
    region = np.zeros(N, dtype=np.int32)
    region[np.random.default_rng(8).choice(N, 9, replace=False)] = 1
    n = 40
    d, terms = P.ref_derivative(beam, L, region, which, stop_rule=O.STOP_DELTA, tol=0.0, max_terms=n)
    assert terms == n
    f = _reference_factor(P, L, which)

    def delta(AW):
        return f[:, None] * AW * region[None, :]

    s0, _, _ = P.source(beam, L)
    # D = c_m g^l mu_s and Delta are both of order c: eps is a relative step
    eps = 1e-6
    Ep = _series_exitance(P, s0, L, n, delta, eps)
    Em = _series_exitance(P, s0, L, n, delta, -eps)
    fd = (Ep - Em) / (2 * eps)
    assert np.max(np.abs(d)) > 0
    assert rel_max(d, fd) < 1e-7
    # the series are converged: 10 fewer terms change neither
    d30, _ = P.ref_derivative(beam, L, region, which, stop_rule=O.STOP_DELTA, tol=0.0, max_terms=n - 10)
    assert rel_max(d30, d) < 1e-12


@pytest.mark.parametrize("which", [0, 1])
def test_reference_derivative_unit_free(which):
    """With C read as the optics' speed of light, the derivative image does not
    depend on the unit of c (the exitance does not): c = 1 (the reference's
    FisherInfo_IO/src/rte.h:15) and c = 2.9979e10 cm/s give the same image at
    n = 1.37.  A source of -1 out2 instead of -C out2 would scale with 1/c."""
    shape, L, g = (4, 3, 4), 2, 0.8
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=4, sigma=0.2)
    region = np.zeros(int(np.prod(shape)), dtype=np.int32)
    region[5:17] = 1
    beam = (0.0, 0.0, 0.1, 1.0)
    imgs = []
    for c in (1.0, 2.9979e10):
        P = O.Problem(shape, bounds, ma, ms, g, n=1.37, c=c)
        imgs.append(P.ref_derivative(beam, L, region, which, stop_rule=O.STOP_DELTA, tol=0.0, max_terms=8)[0])
    assert np.max(np.abs(imgs[0])) > 0
    assert rel_max(imgs[1], imgs[0]) < 1e-12


@pytest.mark.parametrize("which", [0, 1])
def test_reference_derivative_additivity(which):
    shape, L, g = (5, 4, 4), 2, 0.7
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=21, sigma=0.3)
    P = O.Problem(shape, bounds, ma, ms, g, n=1.3)
    beam = (0.0, 0.05, 0.1, 1.0)
    N = int(np.prod(shape))
    rng = np.random.default_rng(5)
    lab = rng.integers(0, 3, N)
    n = 6
    d = {}
    for r in range(3):
        d[r], terms = P.ref_derivative(beam, L, (lab == r).astype(np.int32), which, stop_rule=O.STOP_DELTA,
                                       tol=0.0, max_terms=n)
        assert terms == n
    # the derivative source is linear in the region
    u, _ = P.ref_derivative(beam, L, (lab != 1).astype(np.int32), which, stop_rule=O.STOP_DELTA, tol=0.0,
                            max_terms=n)
    assert rel_max(u, d[0] + d[2]) < 1e-12


@pytest.mark.parametrize("L", [0, 1, 3, 5])
def test_reference_layout_map_preserves_the_radiance(L):
    """doi_amd.rte.to_reference_layout (the map rte_neumann_ref applies on the
    device): the complex coefficients it gives describe the same radiance as the
    real-basis ones, sum_a r_a R_a(u) = sum_lm c_lm Y_lm(u), at random directions,
    with the oracle's complex harmonics (pinned to scipy and closed forms above);
    they satisfy c_l,-m = (-1)^m conj c_lm, and from_reference_layout inverts it."""
    from doi_amd import rte
    K, N = (L + 1) ** 2, 7
    rng = np.random.default_rng(L + 40)
    r = rng.standard_normal((K, N))
    c = rte.to_reference_layout(r, L)
    assert c.shape == (N, K)
    for _ in range(5):
        th, ph = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        Yc = O.ylm_complex(L, th, ph)
        Yr = O.ylm_real(L, th, ph)
        f_real = Yr @ r
        f_cplx = c @ Yc
        assert np.max(np.abs(f_cplx.imag)) < 1e-12
        assert np.max(np.abs(f_cplx.real - f_real)) < 1e-12 * max(1.0, np.max(np.abs(f_real)))
    for l in range(L + 1):
        for m in range(1, l + 1):
            assert np.allclose(c[:, l * l + l - m], (-1) ** m * np.conj(c[:, l * l + l + m]), atol=1e-15)
    assert np.max(np.abs(rte.from_reference_layout(c, L) - r)) < 1e-14
