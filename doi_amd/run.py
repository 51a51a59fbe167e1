"""Run a reference parameter file through librte:

    python -m doi_amd.run file.ini [--precision fp64|mixed|fp32] [--fixed-terms]
                                   [--tol 1e-6] [--out trans.out]

The stop rule is the CPU package's lambda rule (CPU package/src/RTE_uniform.c:95-111)
unless --fixed-terms, which runs nTerms terms as the GPU package's Neumann does
(GPU package/src/neumann.cu:54), or --tol, which stops at a relative change of
the total exitance below tol.  The transmitted image (with the ballistic term)
and the top-face reflectance are written to [Runtime] TransFile and ReflectFile
in the OutputImage format, imaginary part 0 (a continuous-wave solve); with
--out only the transmitted image is written, there.  PhaseFile (the
frequency-domain solve) is not produced.  The photon total (outward and inward
flux through the six faces plus the ballistic transmission,
CPU package/src/Dist.c:538-988) is printed as the reference prints it.
"""
import argparse
import sys

from . import faces, refio, rte

PRECISIONS = {"fp64": rte.PREC_FP64, "mixed": rte.PREC_MIXED, "fp32": rte.PREC_FP32}


def solve_file(path, precision="mixed", fixed_terms=False, tol=None, out=None, ctx=None, quiet=False):
    prob = refio.load(path)
    if prob.modulation != 0.0:
        raise ValueError("a modulated source (frequency-domain solve) is not supported")
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
    try:
        med = ctx.medium(prob.shape, prob.bounds, prob.mu_a, prob.mu_s, prob.g, n=prob.n, c=prob.c)
        try:
            field, img, refl, st = ctx.solve_field(med, prob.beam, params)
        finally:
            med.close()
    finally:
        if own:
            ctx.close()
    st["photons"] = faces.photon_number(field, prob.shape, prob.bounds, prob.L, st["ballistic"], c=prob.c, n=prob.n)
    dest = out or prob.trans_file
    if dest:
        refio.write_image(dest, img)
    if prob.reflect_file and not out:
        refio.write_image(prob.reflect_file, refl)
    if not quiet:
        print("Number of photons total: %f" % st["photons"])
        print("Neumann series: %d" % st["terms"])
        if prob.phase_file:
            print("note: PhaseFile (%s) is not produced" % prob.phase_file, file=sys.stderr)
    return img, st, refl


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("ini")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="mixed")
    ap.add_argument("--fixed-terms", action="store_true")
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    solve_file(a.ini, a.precision, a.fixed_terms, a.tol, a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
