"""Run a reference parameter file through librte:

    python -m doi_amd.run file.ini [--precision fp64|mixed|fp32] [--fixed-terms]
                                   [--tol 1e-6] [--out trans.out]

The stop rule is the CPU package's lambda rule (CPU package/src/RTE_uniform.c:95-111)
unless --fixed-terms, which runs nTerms terms as the GPU package's Neumann does
(GPU package/src/neumann.cu:54), or --tol, which stops at a relative change of
the total exitance below tol.  The transmitted image (with the ballistic term)
and the top-face reflectance are written to [Runtime] TransFile and ReflectFile
in the OutputImage format, the phase image to PhaseFile (degrees,
ComputePhaseImage, CPU package/src/Dist.c:71-127: the argument of the exitance
with the scattered part's quadrature divided by 99 instead of 100); with --out
only the transmitted image is written, there.  Like the reference, the command
line also writes the four side-face images SideFace1(-X).out .. SideFace4(+Y).out
to the working directory (ComputeSideFaceImages, CPU package/src/Dist.c:353-530;
--no-side-faces skips them).  The photon total (outward and inward flux through
the six faces plus the ballistic transmission, CPU package/src/Dist.c:538-988)
is printed as the reference prints it.

A [Source] modulation (angular frequency, rad/s) adds the frequency-domain
solve (rte_solve_fd) of CPU package/src/RTE_uniform.c:120-177: PhaseFile is the
phase of its complex exitance, and TransFile, ReflectFile, the side faces and the
photon total are those of CombineOuts' field (Dist.c:1294-1303), half the continuous-wave
field plus half the real part of the modulated one; imaginary planes are 0.
"""
import argparse
import sys

import numpy as np

from . import faces, refio, rte

PRECISIONS = {"fp64": rte.PREC_FP64, "mixed": rte.PREC_MIXED, "fp32": rte.PREC_FP32}


def ballistic_image(prob, kim=0.0):
    """The ballistic transmission per beam column, fact e^{-sum_k (mu_t + i kim) dz}
    (complex when kim != 0), zero elsewhere; columns and power as the source
    (CPU package/src/Dist.c:1090-1114, Inits.c:56)."""
    nx, ny, nz = prob.shape
    b = prob.bounds
    dx, dy, dz = (b[1] - b[0]) / nx, (b[3] - b[2]) / ny, (b[5] - b[4]) / nz
    xs = np.arange(nx) * dx + b[0] + dx / 2.0
    ys = np.arange(ny) * dy + b[2] + dy / 2.0
    px, py, diam, mag = prob.beam
    inb = np.sqrt((xs[None, :] - px) ** 2 + (ys[:, None] - py) ** 2) <= (diam + 1.0e-16) / 2.0
    tau = (np.asarray(prob.mu_a) + np.asarray(prob.mu_s)).reshape(nz, ny, nx).sum(axis=0) * dz
    col = np.exp(-(tau + 1j * kim * nz * dz)) * (mag / max(int(inb.sum()), 1))
    return np.where(inb, col, 0.0)


def phase_image(exitance, ballistic):
    """ComputePhaseImage (CPU package/src/Dist.c:71-127) in degrees: the scattered
    part enters with the quadrature weights /99 where the exitance has /100."""
    return np.angle(ballistic + (exitance - ballistic) * (100.0 / 99.0)) * 180.0 / np.pi


def solve_file(path, precision="mixed", fixed_terms=False, tol=None, out=None, ctx=None, quiet=False,
               side_faces=False):
    prob = refio.load(path)
    if fixed_terms:
        params = rte.make_params(prob.L, precision=PRECISIONS[precision], stop_rule=rte.STOP_DELTA, tol=0.0,
                                 max_terms=max(prob.n_terms, 1), prop_thresh=prob.prop_thresh)
    elif tol is not None:
        params = rte.make_params(prob.L, precision=PRECISIONS[precision], stop_rule=rte.STOP_DELTA, tol=tol,
                                 max_terms=1000, prop_thresh=prob.prop_thresh)
    else:
        params = rte.make_params(prob.L, precision=PRECISIONS[precision], stop_rule=rte.STOP_LAMBDA,
                                 max_terms=1000, prop_thresh=prob.prop_thresh)
    own = ctx is None
    ctx = ctx or rte.Context(0)
    fd = None
    try:
        med = ctx.medium(prob.shape, prob.bounds, prob.mu_a, prob.mu_s, prob.g, n=prob.n, c=prob.c)
        try:
            field, img, refl, st = ctx.solve_field(med, prob.beam, params)
            if prob.modulation != 0.0:
                fd = ctx.solve_fd([med], [prob.beam], [prob.modulation], params, field=True)
        finally:
            med.close()
    finally:
        if own:
            ctx.close()
    st["photons"] = faces.photon_number(field, prob.shape, prob.bounds, prob.L, st["ballistic"], c=prob.c, n=prob.n)
    ball = ballistic_image(prob)
    phase = phase_image(img, ball)
    if fd is not None:
        # CombineOuts: half the continuous-wave field, half the real part of the modulated one
        img_c, refl_c, st_c, field_c = fd[0][0], fd[1][0], fd[2][0], fd[3]
        ball_c = ballistic_image(prob, prob.modulation * prob.n / prob.c)
        phase = phase_image(img_c, ball_c)
        img = 0.5 * img + 0.5 * img_c.real
        refl = 0.5 * refl + 0.5 * refl_c.real
        st["photons"] = 0.5 * st["photons"] + 0.5 * faces.photon_number(
            np.ascontiguousarray(field_c.real), prob.shape, prob.bounds, prob.L, float(np.sum(ball_c.real)),
            c=prob.c, n=prob.n)
        st["terms_fd"] = st_c["terms"]
        field = 0.5 * field + 0.5 * field_c.real
    st["phase"] = phase
    dest = out or prob.trans_file
    if dest:
        refio.write_image(dest, img)
    if not out:
        if prob.reflect_file:
            refio.write_image(prob.reflect_file, refl)
        if prob.phase_file:
            refio.write_phase_image(prob.phase_file, phase)
        if side_faces:
            faces.write_side_faces(".", faces.side_face_images(field, prob.shape, prob.bounds, prob.L, c=prob.c,
                                                               n=prob.n))
    if not quiet:
        print("Number of photons total: %f" % st["photons"])
        print("Neumann series: %d" % st.get("terms_fd", st["terms"]))
    return img, st, refl


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("ini")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="mixed")
    ap.add_argument("--fixed-terms", action="store_true")
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--no-side-faces", action="store_true",
                    help="do not write SideFace1(-X).out .. SideFace4(+Y).out to the working directory")
    a = ap.parse_args(argv)
    solve_file(a.ini, a.precision, a.fixed_terms, a.tol, a.out, side_faces=not a.no_side_faces)
    return 0


if __name__ == "__main__":
    sys.exit(main())
