"""A reference parameter file run end to end through librte (doi_amd.run):
the transmitted image written in the reference's OutputImage format, and the
reflectance image, equal the reference CPU package's for the same file (FP64
within 1e-9, same lambda-rule term count); the command-line entry with a fixed
term count writes both files and the four side-face files; a modulated source
adds the frequency-domain solve to the files."""
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
        side = {f: refio.read_image(name)[0] for f, name in faces.SIDE_FACE_FILES}
    finally:
        os.chdir(cwd)
    assert rel_max(t, np.array(case["trans_image"])) < 1e-4
    assert rel_max(r, np.array(case["refl_image"])) < 1e-4
    for f, _ in faces.SIDE_FACE_FILES:   # the reference binary's SideFace1..4 files
        assert rel_max(side[f], np.array(case["side_faces"][f])) < 1e-4, f


def test_modulated_parameter_file(ctx, golden_ref, golden_mod, tmp_path):
    """A parameter file with a [Source] modulation: PhaseFile is the phase of
    the frequency-domain exitance (ComputePhaseImage's quadrature), TransFile
    and ReflectFile are CombineOuts' half continuous-wave plus half the real part
    of the modulated solve, all against the oracle's series (FP64 path, 1e-9).
    Without a modulation PhaseFile holds the phase of the real exitance, 0."""
    import os
    cwd = os.getcwd()
    try:
        os.chdir(str(tmp_path))
        for case in golden_mod:
            print("modulated file", case["name"], flush=True)
            (tmp_path / "m.ini").write_text(ini_for(case, mod=case["modulation"]))
            img, st, refl = run.solve_file("m.ini", precision="fp64", ctx=ctx, quiet=True)
            t, ti = refio.read_image("trans.out")
            r, ri = refio.read_image("refl.out")
            ph = refio.read_phase_image("phase.out")
            b = case["beam"]
            beam = (b["pos_x"], b["pos_y"], b["diam"], 1.0)
            P = O.Problem(tuple(case["shape"]), tuple(case["bounds"]), case["mua"], case["mus"], case["g"],
                          n=case["n"])
            dc = P.solve(beam, case["L"], prop_thresh=case["prop_thresh"], reflectance=True)
            ac = P.solve_fd(beam, case["L"], case["modulation"], prop_thresh=case["prop_thresh"])
            assert st["terms"] == dc["terms"] and st["terms_fd"] == ac["terms"]
            assert rel_max(t, 0.5 * dc["exitance"] + 0.5 * ac["exitance"].real) < 1e-9
            assert rel_max(r, 0.5 * dc["reflectance"] + 0.5 * ac["reflectance"].real) < 1e-9
            assert not np.any(ti) and not np.any(ri)
            prob = refio.load("m.ini")
            want = run.phase_image(ac["exitance"], run.ballistic_image(prob, case["modulation"] * case["n"] / 2.9979e10))
            assert np.max(np.abs(ph - want)) < 1e-6
            assert np.min(ph) < -1.0                      # degrees of delay
        (tmp_path / "c.ini").write_text(ini_for(golden_ref[0]))
        run.solve_file("c.ini", precision="fp64", ctx=ctx, quiet=True)
        assert not np.any(refio.read_phase_image("phase.out"))
    finally:
        os.chdir(cwd)


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
