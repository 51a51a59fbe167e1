"""Fisher information of a small absorbing region against its depth and size,
the reference's FIM scan (FisherInfo_IO/src/rte_fim.cu:80-379):

    F = sum_j (dg_j)^2 / g_j          over the pixels with g_j != 0   (rte_fim.cu:318-332)

g is the top-layer image of the forward field (generate_ref_image,
FisherInfo_IO/src/dist.cu:229-269; `snr.fim_top_image` here), and dg the same
image of the reference's forward-mode derivative with respect to mu_a of the
voxels whose centres lie within `radius` of voxel (nx/2, ny/2, z), inclusive
(rte_fim.cu:256-273 zeroes the field where the squared distance exceeds
radius^2, :265): one Neumann series (rte_neumann) from the forward field restricted
to that region, times -C (rte_fim.cu:274-277), C the vacuum speed of light
(FisherInfo_IO/src/rte.h:15; the optics' c here).  The reference's radii are (i + 1) fov / (2 nx),
i = 0..4, fov = xmax - xmin (rte_fim.cu:209-216), its depths z = 1 .. nz - 3
(the loops at :252-253, commented out: the program runs z = initz = 1 and the
smallest radius, :254-255, which holds the centre voxel alone).  It sets every
tissue to mu_a 0.01, mu_s 1.0 (:226-229); here the maps are the caller's.

Differences on purpose: the voxel centres are the grid's own (the reference
derives all three axes from the x extent with zmin = 0, :173-179, identical on
a cube [-a, a]^2 x [0, 2a]); the scan returns every (depth, radius) it is given
instead of printing one line each.
"""
import numpy as np

from . import rte
from .snr import fim_top_image, voxel_centres


def reference_radii(shape, bounds, count=5):
    """(i + 1) fov / (2 nx), i = 0 .. count - 1 (rte_fim.cu:209-216)."""
    fov = bounds[1] - bounds[0]
    return [(i + 1) * fov / (2.0 * shape[0]) for i in range(count)]


def region_mask(shape, bounds, centre, radius):
    """Voxels (x fastest, flat) whose centre is within radius of the centre of
    voxel `centre` = (i, j, k), the boundary included (rte_fim.cu:265)."""
    xs, ys, zs = voxel_centres(shape, bounds)
    i, j, k = centre
    d2 = ((xs[None, None, :] - xs[i]) ** 2 + (ys[None, :, None] - ys[j]) ** 2 + (zs[:, None, None] - zs[k]) ** 2)
    return (d2 <= radius * radius).ravel()


def fisher_top(g, dg):
    """sum_j dg_j^2 / g_j over the pixels with g_j != 0 (rte_fim.cu:318-324)."""
    g = np.ravel(g)
    dg = np.ravel(dg)
    ok = g != 0
    return float(np.sum(dg[ok] ** 2 / g[ok]))


def fim_scan(ctx, shape, bounds, mua, mus, g, beam, params, n=1.0, c=2.9979e10, depths=(1,), radii=None):
    """F for every (depth index z, radius) pair, the region centred on voxel
    (nx/2, ny/2, z).  radii defaults to the reference's first radius.  Returns
    dict(depths, radii, fisher [len(depths), len(radii)], g, dg {(z, r): image})."""
    nx, ny, nz = shape
    if radii is None:
        radii = reference_radii(shape, bounds)[:1]
    for z in depths:
        if not 0 <= z < nz:
            raise ValueError("depth index %d outside the grid" % z)
    m = ctx.medium(shape, bounds, mua, mus, g, n=n, c=c)
    try:
        field, _, _, _ = ctx.solve_field(m, beam, params)
        g0 = fim_top_image(field, shape, bounds, params.L, c=c, n=n, mag=beam[3])
        F = np.zeros((len(depths), len(radii)))
        dgs = {}
        for a, z in enumerate(depths):
            for b, rad in enumerate(radii):
                inside = region_mask(shape, bounds, (nx // 2, ny // 2, z), rad)
                src = np.where(inside[None, :], -c * field, 0.0)
                dfield, _, _, _ = ctx.neumann(m, src, params)
                dg = fim_top_image(dfield, shape, bounds, params.L, c=c, n=n, mag=beam[3])
                dgs[(z, rad)] = dg
                F[a, b] = fisher_top(g0, dg)
    finally:
        m.close()
    return dict(depths=list(depths), radii=list(radii), fisher=F, g=g0, dg=dgs)


def main(argv=None):
    """python -m doi_amd.fim file.ini [--all]: the reference's FIM scan of a
    parameter file (mu maps as loaded; --all scans z = 1 .. nz - 3 and the five
    radii, as the reference's commented-out loops would)."""
    import argparse
    from . import refio
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("ini")
    ap.add_argument("--all", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed")
    a = ap.parse_args(argv)
    prob = refio.load(a.ini)
    params = rte.make_params(prob.L, precision=rte.PREC_FP64 if a.precision == "fp64" else rte.PREC_MIXED,
                             stop_rule=rte.STOP_LAMBDA, max_terms=1000, prop_thresh=prob.prop_thresh)
    depths = list(range(1, prob.shape[2] - 2)) if a.all else [1]
    radii = reference_radii(prob.shape, prob.bounds) if a.all else None
    ctx = rte.Context(0)
    try:
        res = fim_scan(ctx, prob.shape, prob.bounds, prob.mu_a, prob.mu_s, prob.g, prob.beam, params, n=prob.n,
                       c=prob.c, depths=depths, radii=radii)
    finally:
        ctx.close()
    for i, z in enumerate(res["depths"]):
        for j, r in enumerate(res["radii"]):
            print("%e is fisher abs for %d as z index and %f as radius" % (res["fisher"][i, j], z, r))
    return 0


__all__ = ["fim_scan", "fisher_top", "reference_radii", "region_mask"]

if __name__ == "__main__":
    import sys
    sys.exit(main())
