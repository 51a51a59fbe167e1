"""A reference parameter file run end to end through librte (doi_amd.run):
the transmitted image written in the reference's OutputImage format, and the
reflectance image, equal the reference CPU package's for the same file (FP64
within 1e-9, same lambda-rule term count); the command-line entry with a fixed
term count writes both files."""
import numpy as np
import pytest

from doi_amd import faces, refio, rte, run
from oracle import oracle as O
from tests.helpers import bounds_for, hetero_maps, rel_max
from tests.test_refio import ini_for

pytestmark = pytest.mark.gpu


def test_parameter_file_end_to_end(ctx, golden_ref, tmp_path):
    for case in (golden_ref[0], golden_ref[6]):
        p = tmp_path / (case["name"] + ".ini")
        p.write_text(ini_for(case))
        out = tmp_path / (case["name"] + ".trans")
        img, st, refl = run.solve_file(str(p), precision="fp64", out=str(out), ctx=ctx, quiet=True)
        real, imag = refio.read_image(str(out))
        assert st["terms"] == case["terms"]
        assert real.shape == (case["shape"][1], case["shape"][0]) and not np.any(imag)
        assert rel_max(real, np.array(case["trans_image"])) < 1e-9, case["name"]
        assert rel_max(refl, np.array(case["refl_image"])) < 1e-9, case["name"]
        assert abs(st["photons"] - case["photons_printed"]) <= 5.0e-7 + 1e-9, case["name"]
    # the command line: nTerms = 30 terms, mixed precision, against the lambda-rule FP64 image
    case = golden_ref[0]
    p = tmp_path / "cli.ini"
    p.write_text(ini_for(case))
    out = tmp_path / "cli.trans"
    assert run.main([str(p), "--fixed-terms", "--out", str(out)]) == 0
    real, _ = refio.read_image(str(out))
    assert rel_max(real, np.array(case["trans_image"])) < 1e-4
    # without --out: [Runtime] TransFile and ReflectFile, relative to the working directory
    import os
    cwd = os.getcwd()
    try:
        os.chdir(str(tmp_path))
        assert run.main([str(p), "--fixed-terms"]) == 0
        t, _ = refio.read_image("trans.out")
        r, _ = refio.read_image("refl.out")
    finally:
        os.chdir(cwd)
    assert rel_max(t, np.array(case["trans_image"])) < 1e-4
    assert rel_max(r, np.array(case["refl_image"])) < 1e-4


def test_final_field_against_oracle(ctx):
    """rte_solve_field: the final radiance field (reference units) of a
    heterogeneous solve against the oracle's (FP64 1e-10, mixed 1e-5 of the
    largest coefficient), and the face fluxes computed from it."""
    shape, L = (6, 5, 4), 3
    bounds = bounds_for(shape)
    ma, ms = hetero_maps(shape, seed=71)
    beam = (0.05, 0.0, 0.1, 1.0)
    ref = O.Problem(shape, bounds, ma, ms, 0.8, n=1.2).solve(beam, L, stop_rule=O.STOP_DELTA, tol=1e-7,
                                                           max_terms=30, field=True)
    m = ctx.medium(shape, bounds, ma, ms, 0.8, n=1.2)
    for prec, tol in ((rte.PREC_FP64, 1e-10), (rte.PREC_MIXED, 1e-5)):
        p = rte.make_params(L, precision=prec, stop_rule=rte.STOP_DELTA, tol=1e-7, max_terms=30)
        field, img, refl, st = ctx.solve_field(m, beam, p)
        assert st["terms"] == ref["terms"]
        assert np.max(np.abs(field - ref["w"])) < tol * np.max(np.abs(ref["w"]))
        assert rel_max(img, ref["exitance"]) < tol
        t_gpu = faces.photon_number(field, shape, bounds, L, st["ballistic"], n=1.2)
        t_orc = faces.photon_number(ref["w"], shape, bounds, L, ref["ballistic_total"], n=1.2)
        assert abs(t_gpu - t_orc) < 10 * tol * abs(t_orc)
    m.close()
