"""Ideal-observer detectability of a small absorbing inclusion, the reference's
SNR study (FisherInfo_IO/src/rte_snr.cu:203-262): for each inclusion depth, the
mean images with the signal present (g1) and absent (g0) from two solves, and

    SNR^2 = sum_j (g1_j - g0_j)^2 / g0_j                       (rte_snr.cu:244-250)

the ideal observer's detectability of a known signal under Poisson noise.  The
inclusion is tissue 2, the voxels whose centres lie within `radius` of the
inclusion centre (strictly, :219); its mu_a is mu_absent + delta when present
and mu_absent when absent, mu_s = mus_inc (:227-231).  The reference scans the
centre (nX/2, nY/2, z) for z = 2 .. 15 (:213-216) with radius 3 dx / 2
(rad_sig[2], :196-201).  Its images are generate_ref_image's (FisherInfo_IO/src/
dist.cu:229-269), restated here as `fim_top_image`.

Where this differs from the reference, on purpose: the voxel centres are the
grid's own (the reference derives x, y and z from the x extent and assumes
zmin = 0, rte_snr.cu:168-176, identical for a cube [-a, a]^2 x [0, 2a]); the
depth loop advances by one (its increment `loc_indz = loc_indz++` is undefined
behaviour in C) and stops at nz; pixels with g0 = 0 are left out of the sum.

`linearized_snr2` gives the small-signal limit delta^2 F from the exact Fisher
information of the inclusion's mu_a (rte_fisher, transmitted image), which the
two-solve SNR^2 approaches as delta -> 0.
"""
import math

import numpy as np

IMAGES = ("fim_top", "reflectance", "transmission")


def voxel_centres(shape, bounds):
    nx, ny, nz = shape
    b = bounds
    dx, dy, dz = (b[1] - b[0]) / nx, (b[3] - b[2]) / ny, (b[5] - b[4]) / nz
    return (b[0] + (np.arange(nx) + 0.5) * dx, b[2] + (np.arange(ny) + 0.5) * dy,
            b[4] + (np.arange(nz) + 0.5) * dz)


def inclusion_mask(shape, bounds, centre, radius):
    """Voxels (x fastest, flat) whose centre is strictly within radius of the
    centre of voxel `centre` = (i, j, k)."""
    xs, ys, zs = voxel_centres(shape, bounds)
    i, j, k = centre
    d2 = ((xs[None, None, :] - xs[i]) ** 2 + (ys[None, :, None] - ys[j]) ** 2 + (zs[:, None, None] - zs[k]) ** 2)
    return (d2 < radius * radius).ravel()


def fim_top_weights(L):
    """Q_l = sum_{p=0}^{99} P_l(x_p) x_p / 99, x_p = p / 99 (FisherInfo_IO/src/dist.cu:242-261)."""
    x = np.arange(100) / 99.0
    out = np.zeros(L + 1)
    for l in range(L + 1):
        c = np.zeros(l + 1)
        c[l] = 1.0
        out[l] = np.sum(np.polynomial.legendre.legval(x, c) * x) / 99.0
    return out


def fim_top_image(field, shape, bounds, L, c=2.9979e10, n=1.0, mag=1.0):
    """generate_ref_image: sum_l 2 pi dx dy c_m / sqrt(4 pi) sqrt(2l+1) Q_l w_l0 on
    the top layer (z index 0), times the beam magnitude (FisherInfo_IO/src/
    dist.cu:250-263); field is rte_solve_field's (reference units)."""
    nx, ny, nz = shape
    dx, dy = (bounds[1] - bounds[0]) / nx, (bounds[3] - bounds[2]) / ny
    K = (L + 1) ** 2
    f = np.asarray(field).reshape(K, nz, ny, nx)
    Q = fim_top_weights(L)
    img = np.zeros((ny, nx))
    for l in range(L + 1):
        img += 2 * math.pi * dx * dy * (c / n) / math.sqrt(4 * math.pi) * math.sqrt(2 * l + 1) * Q[l] * \
            f[l * l + l, 0]
    return img * mag


def snr2(g1, g0):
    """sum_j (g1_j - g0_j)^2 / g0_j over the pixels with g0_j > 0."""
    g1 = np.ravel(g1)
    g0 = np.ravel(g0)
    ok = g0 > 0
    return float(np.sum((g1[ok] - g0[ok]) ** 2 / g0[ok]))


def _images(ctx, media, beam, params, image, shape, bounds, n, c):
    if image == "fim_top":
        out = []
        for m in media:
            field, _, _, _ = ctx.solve_field(m, beam, params)
            out.append(fim_top_image(field, shape, bounds, params.L, c=c, n=n, mag=beam[3]))
        return out
    img, refl, _ = ctx.solve_images(media, [beam] * len(media), params)
    return list(refl if image == "reflectance" else img)


def depth_scan(ctx, shape, bounds, mua_bg, mus_bg, g, beam, params, n=1.0, c=2.9979e10, radius=None,
               depths=None, mua_absent=0.25, delta=0.1, mus_inc=1.0, image="fim_top"):
    """SNR^2 of the inclusion at the centre (nx/2, ny/2, z) for each z in depths
    (default 2 .. min(16, nz) - 1).  Returns dict(depths, snr2, g1, g0)."""
    if image not in IMAGES:
        raise ValueError("image must be one of %s" % (IMAGES,))
    nx, ny, nz = shape
    if radius is None:
        radius = 3.0 * (bounds[1] - bounds[0]) / (2.0 * nx)
    if depths is None:
        depths = list(range(2, min(16, nz)))
    N = nx * ny * nz
    res = dict(depths=list(depths), snr2=[], g1=[], g0=[])
    for z in depths:
        if not 0 <= z < nz:
            raise ValueError("depth index %d outside the grid" % z)
        mask = inclusion_mask(shape, bounds, (nx // 2, ny // 2, z), radius)
        media = []
        for mu_inc in (mua_absent + delta, mua_absent):
            ma = np.where(mask, mu_inc, np.broadcast_to(mua_bg, (N,)))
            ms = np.where(mask, mus_inc, np.broadcast_to(mus_bg, (N,)))
            media.append(ctx.medium(shape, bounds, ma, ms, g, n=n, c=c))
        try:
            g1, g0 = _images(ctx, media, beam, params, image, shape, bounds, n, c)
        finally:
            for m in media:
                m.close()
        res["g1"].append(g1)
        res["g0"].append(g0)
        res["snr2"].append(snr2(g1, g0))
    return res


def linearized_snr2(ctx, shape, bounds, mua, mus, g, beam, params, mask, delta, n=1.0, c=2.9979e10):
    """delta^2 F, F = sum_j (dE_j/dmu)^2 / E_j: the Fisher information of one
    parameter, the mu_a of the voxels in mask, on the transmitted image."""
    m = ctx.medium(shape, bounds, mua, mus, g, n=n, c=c)
    try:
        part = np.where(np.ravel(mask), 0, -1).astype(np.int32)
        F, _, _, _, _ = ctx.fisher(m, beam, params, part, 1)
    finally:
        m.close()
    return float(delta * delta * F[0, 0])


__all__ = ["depth_scan", "linearized_snr2", "fim_top_image", "fim_top_weights", "inclusion_mask", "snr2",
           "voxel_centres", "IMAGES"]
