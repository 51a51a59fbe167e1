"""A reference parameter file run end to end through librte (doi_amd.run):
the transmitted image written in the reference's OutputImage format, and the
reflectance image, equal the reference CPU package's for the same file (FP64
within 1e-9, same lambda-rule term count); the command-line entry with a fixed
term count writes both files."""
import numpy as np
import pytest

from doi_amd import refio, run
from tests.helpers import rel_max
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
