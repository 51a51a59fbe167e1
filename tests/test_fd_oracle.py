"""Frequency-domain (modulated source) restatement in the oracle, pinned on the CPU.

The reference's modulated branch (CPU package/src/RTE_uniform.c:120-177) does not
compute the frequency-domain series it sets out to: its second source is built
before mu_a turns complex and is normalised by the beam columns twice (Dist.c:
1113-1114 overwrites source->mag), its series adds into the continuous-wave field
`out` and stops on it, and `out2` ends as one complex propagation of that source.
Composed in exactly that order, the oracle's complex pieces (propagation with the
complex pair factor and self block, complex ballistic term) reproduce every
output of the reference binary run with a modulation: the transmitted and
reflected images (both planes), the phase image (through doi_amd.run's
ballistic_image and phase_image) and the term counts.  The other
tests pin what that composition leaves open: the complex source against the
reference's closed form (Dist.c:1097-1103) with a complex mu_t, the pure
absorber, omega = 0, conjugate symmetry (delta rule), the heterogeneous code path and the
complex stop-rule sums against a direct projection on the complex harmonics."""
from types import SimpleNamespace

import numpy as np

from doi_amd import run
from oracle import oracle as O

C_LIGHT = 2.9979e10


def _stop(fi, n, sw, so):
    """StopIteration's lambda (CPU package/src/Kernels_selfsub.c:1208-1233)."""
    if n <= 1:
        return False, fi
    if so == 0.0:
        return True, fi
    gam = sw / so
    hi = 1.0 - gam * (1.0 / (1.0 - gam / fi))
    if hi > 1.0 or hi < 0.0:
        hi = 0.0
    return hi > 0.9999, gam


def _functional(P, w, L, which):
    """The real exitance / reflectance functional applied to a complex field."""
    f = P.exitance if which == "trans" else P.reflectance
    return f(np.ascontiguousarray(w.real), L) + 1j * f(np.ascontiguousarray(w.imag), L)


def _reference_modulated(case):
    """The reference main's modulated path, step by step, on the oracle's pieces."""
    nx, ny, nz = case["shape"]
    b = case["bounds"]
    L, g, om, pt = case["L"], case["g"], case["modulation"], case["prop_thresh"]
    P = O.Problem((nx, ny, nz), tuple(b), case["mua"], case["mus"], g, n=case["n"])
    beam = (case["beam"]["pos_x"], case["beam"]["pos_y"], case["beam"]["diam"], 1.0)
    K = (L + 1) ** 2
    cw = P.solve(beam, L, stop_rule=O.STOP_LAMBDA, prop_thresh=pt, field=True)
    s0, _, cols = P.source(beam, L)
    src2 = (s0 / cols).astype(np.complex128)            # normalised twice
    out = cw["w"].astype(np.complex128)                 # the CW field after its final transform
    W = src2.copy()
    Lloop = L if g != 0.0 else 0
    Kl = (Lloop + 1) ** 2
    fi, n = 1.0, 1
    while True:
        stop, fi = _stop(fi, n, P.bottom_sum_c(W, L), P.bottom_sum_c(out, L))
        if stop:
            break
        AW = np.zeros_like(W)
        AW[:Kl] = P.propagate_c(W[:Kl], Lloop, om, prop_thresh=pt)
        W = P.scatter(AW.real, L) + 1j * P.scatter(AW.imag, L)
        out += W
        n += 1
    out2 = P.propagate_c(src2, L, om, prop_thresh=pt)
    # ballistic with source->mag = 1/cols^2 and the complex mu_t (zmin = 0 in
    # every case, so the reference's z = zmax is the depth); the phase image
    # through the CLI's own quadrature (doi_amd.run.phase_image)
    N = nx * ny * nz
    prob = SimpleNamespace(shape=(nx, ny, nz), bounds=tuple(b), beam=beam, mu_a=np.full(N, case["mua"]),
                           mu_s=np.full(N, case["mus"]))
    assert b[4] == 0.0
    ball = (run.ballistic_image(prob, om * case["n"] / C_LIGHT) / cols).ravel()
    e_out2 = _functional(P, out2, L, "trans")
    phase = run.phase_image(ball + e_out2, ball)
    comb = 0.5 * out + 0.5 * out2.real                  # CombineOuts (Dist.c:1294-1303)
    trans = ball + _functional(P, comb, L, "trans")
    refl = _functional(P, comb, L, "refl")
    assert K == comb.shape[0]
    return dict(terms=[cw["terms"], n, n], trans=trans.reshape(ny, nx), refl=refl.reshape(ny, nx),
                phase=phase.reshape(ny, nx))


def test_reference_modulated_outputs_from_oracle_pieces(golden_mod):
    for case in golden_mod:
        r = _reference_modulated(case)
        assert r["terms"] == case["terms_printed"], case["name"]
        tr = np.array(case["trans_image"]) + 1j * np.array(case["trans_image_imag"])
        rf = np.array(case["refl_image"]) + 1j * np.array(case["refl_image_imag"])
        assert np.max(np.abs(r["trans"] - tr)) <= 1e-12 * np.max(np.abs(tr)), case["name"]
        assert np.max(np.abs(r["refl"] - rf)) <= 1e-12 * np.max(np.abs(rf)), case["name"]
        assert np.max(np.abs(r["phase"] - np.array(case["phase_image"]))) < 1e-9, case["name"]
        # the imaginary parts are not negligible: the complex pieces are exercised
        assert np.max(np.abs(tr.imag)) > 1e-3 * np.max(np.abs(tr.real))


def test_complex_source_matches_reference_closed_form():
    """GenerateSourceBeam's m = 0 entries with a complex mu_t (Dist.c:1097-1103):
    g^l (mu_s/mu_t) sqrt((2l+1)/4pi) (e^{-mu_t z} - e^{-mu_t (z+dz)}) / dV on the
    beam columns; the oracle builds them as column products."""
    shape, b = (5, 4, 6), (-0.25, 0.25, -0.2, 0.2, 0.0, 0.6)
    L, g, n, om = 3, 0.7, 1.4, 4.0e10
    P = O.Problem(shape, b, 0.03, 1.7, g, n=n)
    beam = (0.0, 0.0, 0.25, 1.0)
    s0, ball, cols = P.source_c(beam, L, om)
    nx, ny, nz = shape
    dx, dy, dz = 0.1, 0.1, 0.1
    mu = 0.03 + 1.7 + 1j * om * n / C_LIGHT
    xs = b[0] + (np.arange(nx) + 0.5) * dx
    ys = b[2] + (np.arange(ny) + 0.5) * dy
    inb = (np.sqrt(xs[None, :] ** 2 + ys[:, None] ** 2) <= (beam[2] + 1e-16) / 2).ravel()
    assert cols == inb.sum() > 0
    zs = b[4] + (np.arange(nz) + 0.5) * dz
    for l in range(L + 1):
        col = g ** l * (1.7 / mu) * np.sqrt((2 * l + 1) / (4 * np.pi)) * \
            (np.exp(-mu * (zs - dz / 2)) - np.exp(-mu * (zs + dz / 2))) / (dx * dy * dz) / cols
        want = np.where(inb[None, :], col[:, None], 0.0).ravel()
        assert np.max(np.abs(s0[l * l + l] - want)) < 1e-12 * np.max(np.abs(want))
        for m in range(-l, l + 1):
            if m:
                assert not np.any(s0[l * l + l + m])
    assert np.allclose(ball[inb], np.exp(-mu * 0.6) / cols, rtol=1e-13, atol=0)
    assert not np.any(ball[~inb])


def test_pure_absorber_and_zero_frequency():
    shape, b = (4, 4, 5), (-0.2, 0.2, -0.2, 0.2, 0.0, 0.5)
    beam = (0.0, 0.0, 0.2, 1.0)
    om = 3.0e10
    # no scattering: one term, the exitance is the ballistic term e^{-(mu_a + i k) depth}
    Pa = O.Problem(shape, b, 0.4, 0.0, 0.8, n=1.3)
    ra = Pa.solve_fd(beam, 2, om)
    e = ra["exitance"][ra["exitance"] != 0]
    assert e.size == 4
    assert np.allclose(e, 0.25 * np.exp(-(0.4 + 1j * om * 1.3 / C_LIGHT) * 0.5), rtol=1e-14, atol=0)
    # omega = 0 is the continuous-wave solve, bit for bit, both stop rules
    rng = np.random.default_rng(3)
    ma = 0.02 * np.exp(0.3 * rng.standard_normal(80))
    ms = 1.1 * np.exp(0.3 * rng.standard_normal(80))
    P = O.Problem(shape, b, ma, ms, 0.8, n=1.2)
    for rule in (O.STOP_LAMBDA, O.STOP_DELTA):
        r0 = P.solve(beam, 3, stop_rule=rule, tol=1e-8, field=True)
        rf = P.solve_fd(beam, 3, 0.0, stop_rule=rule, tol=1e-8, field=True)
        assert rf["terms"] == r0["terms"]
        assert np.array_equal(rf["exitance"].real, r0["exitance"]) and not np.any(rf["exitance"].imag)
        assert np.array_equal(rf["w"].real, r0["w"]) and not np.any(rf["w"].imag)
    # conjugate symmetry in omega, under the delta rule (the lambda rule's real
    # parts of complex-basis coefficients take, for odd m, the imaginary
    # sin-type coefficients, which change sign with omega)
    rp = P.solve_fd(beam, 3, om, stop_rule=O.STOP_DELTA, tol=1e-8)
    rm = P.solve_fd(beam, 3, -om, stop_rule=O.STOP_DELTA, tol=1e-8)
    assert rp["terms"] == rm["terms"]
    assert np.max(np.abs(rp["exitance"] - np.conj(rm["exitance"]))) < 1e-14 * np.max(np.abs(rp["exitance"]))
    assert np.max(np.abs(rp["exitance"].imag)) > 1e-3 * np.max(np.abs(rp["exitance"].real))


def test_heterogeneous_complex_path_reduces_to_uniform():
    """A medium that differs from a uniform one only in a voxel the field does
    not reach goes through the per-voxel moment series and the marched path:
    the same complex A as the uniform path (direct quadrature, tau = mu r)."""
    shape, b = (5, 4, 4), (-0.25, 0.25, -0.2, 0.2, 0.0, 0.4)
    L, om = 3, 5.0e10
    N = 80
    Pu = O.Problem(shape, b, 0.05, 1.5, 0.8, n=1.3)
    ma = np.full(N, 0.05)
    ma[-1] += 1e-13
    Ph = O.Problem(shape, b, ma, 1.5, 0.8, n=1.3)
    rng = np.random.default_rng(5)
    W = (rng.standard_normal(((L + 1) ** 2, N)) + 1j * rng.standard_normal(((L + 1) ** 2, N)))
    W[:, -1] = 0.0
    Au = Pu.propagate_c(W, L, om)
    Ah = Ph.propagate_c(W, L, om)
    assert np.max(np.abs(Au - Ah)) < 1e-11 * np.max(np.abs(Au))
    # the complex self block: direct quadrature and moment series agree
    mu = 1.55 + 1j * om * 1.3 / C_LIGHT
    S0 = O.self_block_c(L, 0.1, 0.1, 0.1, mu, 1.0, 0)
    S1 = O.self_block_c(L, 0.1, 0.1, 0.1, mu, 1.0, 1)
    assert np.max(np.abs(S0 - S1)) < 1e-12 * np.max(np.abs(S0))
    # and the real self block at a real mu_t
    assert np.array_equal(O.self_block_c(L, 0.1, 0.1, 0.1, 1.55 + 0j, 1.0, 0), O.self_block(L, 0.1, 0.1, 0.1, 1.55))


def test_complex_stop_sums_against_projection():
    """sum over l, m of Re c_lm (StopIteration on a complex field) where the
    complex-basis coefficients c_lm are projected directly from the radiance
    sum_a r_a R_a(omega) on a Gauss x uniform grid (exact to degree 2L)."""
    L = 3
    K = (L + 1) ** 2
    shape, b = (1, 1, 1), (0.0, 0.1, 0.0, 0.1, 0.0, 0.1)
    P = O.Problem(shape, b, 0.1, 1.0, 0.8)
    rng = np.random.default_rng(7)
    r = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    xg, wg = np.polynomial.legendre.leggauss(L + 2)
    nph = 2 * L + 2
    total = 0.0
    for x, wx in zip(xg, wg):
        for j in range(nph):
            th, ph = np.arccos(x), 2 * np.pi * j / nph
            f = np.dot(r, O.ylm_real(L, th, ph))
            total += wx * (2 * np.pi / nph) * np.sum((f * np.conj(O.ylm_complex(L, th, ph))).real)
    got = P.bottom_sum_c(r.reshape(K, 1), L)
    assert abs(got - total) < 1e-12 * max(1.0, abs(total))
    # for a real radiance it is the continuous-wave sum
    assert abs(P.bottom_sum_c(r.real.reshape(K, 1).astype(complex), L) - P.bottom_sum(r.real.reshape(K, 1), L)) < 1e-14
