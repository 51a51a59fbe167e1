"""The SNR study's host pieces (doi_amd.snr), CPU only: the top-image
quadrature of FisherInfo_IO/src/dist.cu:242-261 with the reference's own
associated Legendre function (the oracle's restatement of CPU package/src/
SphHarm.c) and against the exact integrals it approximates; the inclusion
geometry against the reference's coordinate formulas (FisherInfo_IO/src/
rte_snr.cu:168-176, 196-201, 213-220) on a cube; the SNR^2 sum."""
import numpy as np

from doi_amd import snr
from oracle import oracle as O


def test_fim_top_weights():
    L = 6
    Q = snr.fim_top_weights(L)
    for l in range(L + 1):
        ref = sum(O.lib().orc_assoc_legendre(p / 99.0, l, 0) * (p / 99.0) / 99.0 for p in range(100))
        assert abs(Q[l] - ref) < 1e-14
    exact = [1 / 2, 1 / 3, 1 / 8, 0.0, -1 / 48, 0.0, 1 / 128]
    for l in range(L + 1):
        assert abs(Q[l] - exact[l]) < 0.011      # a right-endpoint Riemann sum with 100 points / 99
    # it is not the reflectance quadrature -int_{-1}^0 P_l u du (the reference's other
    # readout, upward directions): the odd orders enter with the opposite sign
    R = -O.quad_weights_refl(L)
    assert abs(Q[0] - R[0]) < 1e-4 and Q[1] * R[1] < 0 and abs(Q[1] + R[1]) < 0.011


def test_inclusion_geometry_matches_reference_on_a_cube():
    n = 16
    a = 0.8
    shape, bounds = (n, n, n), (-a, a, -a, a, 0.0, 2 * a)
    fov = bounds[1] - bounds[0]
    i = np.arange(n)
    xr = i * fov / n - fov / 2 + fov / (2 * n)
    zr = i * fov / n + fov / (2 * n)
    xs, ys, zs = snr.voxel_centres(shape, bounds)
    assert np.allclose(xs, xr, atol=1e-15) and np.allclose(ys, xr, atol=1e-15) and np.allclose(zs, zr, atol=1e-15)
    rad = 3 * fov / (2 * n)                      # rad_sig[2]
    for z in (2, 7, 15):
        m = snr.inclusion_mask(shape, bounds, (n // 2, n // 2, z), rad).reshape(n, n, n)
        want = ((xr[None, None, :] - xr[n // 2]) ** 2 + (xr[None, :, None] - xr[n // 2]) ** 2 +
                (zr[:, None, None] - zr[z]) ** 2) < rad * rad
        assert np.array_equal(m, want)
        assert m.sum() == (19 if z < 15 else 14)   # the sphere of 1.5 voxels, cut by the bottom face at z = 15


def test_snr2_sum():
    g0 = np.array([[1.0, 2.0], [0.0, 4.0]])
    g1 = np.array([[1.5, 2.0], [3.0, 3.0]])
    assert snr.snr2(g1, g0) == 0.25 + 0.0 + 0.25     # the g0 = 0 pixel is left out
