"""The C-ABI library loads and exports every entry point include/rte.h
declares; host-side argument checks that need no device; seeded workloads.
CPU only (no compute call is made)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np

from doi_amd import media, rte

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rte_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    declared = _declared("rte.h")
    assert declared, "no declarations parsed"
    assert sorted(rte.SIGNATURES) == declared


def test_library_exports_every_declared_symbol():
    lib = rte.lib()
    for name in _declared("rte.h"):
        assert hasattr(lib, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", rte.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rte_\w+)", nm))
    assert set(_declared("rte.h")) <= exported


def test_library_targets_gfx950_only():
    data = open(rte.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
    for other in (b"--gfx90a", b"--gfx942", b"--gfx1100"):
        assert other not in data


def test_argument_errors_without_device():
    lib = rte.lib()
    assert lib.rte_ctx_create(0, None) == rte.E_ARG
    assert b"NULL" in lib.rte_last_error()
    p = rte.make_params(5)
    assert lib.rte_solve_media(None, 1, None, None, C.byref(p), None, None) == rte.E_ARG
    assert lib.rte_ctx_profile(None, 1) == rte.E_ARG
    assert lib.rte_medium_reset_cache(None) == rte.E_ARG
    assert lib.rte_version() >= 1
    bad = rte.make_params(rte.MAX_L + 1)
    assert rte.footprint((8, 8, 8), (0, 1, 0, 1, 0, 1), bad) == -1


def test_footprint_scales_with_path_table():
    p = rte.make_params(5, precision=rte.PREC_MIXED)
    b = media.cube_bounds(32)
    cached = rte.footprint((32, 32, 32), b, p, rte.PATH_CACHED)
    march = rte.footprint((32, 32, 32), b, p, rte.PATH_MARCH)
    assert cached - march == 32768 * 32768 * 4      # the e^{-tau}/r^2 table
    p64 = rte.make_params(5, precision=rte.PREC_FP64)
    assert rte.footprint((32, 32, 32), b, p64, rte.PATH_MARCH) > march


def test_media_are_seeded_and_shaped():
    a1, s1 = media.heterogeneous(16, 2)
    a2, s2 = media.heterogeneous(16, 2)
    a3, _ = media.heterogeneous(16, 3)
    assert np.array_equal(a1, a2) and np.array_equal(s1, s2) and not np.array_equal(a1, a3)
    assert a1.shape == (4096,) and np.all(a1 > 0) and np.all(s1 > 0)
    # log-normal factors with sigma 0.2 around the background
    z = np.log(s1 / 1.0)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 0.2) < 0.02
    c5 = media.lognormal_media_with_inclusions(2, seed=5, n=12)
    assert len(c5) == 2 and c5[0][0].shape == (1728,)
    beams = media.top_face_beams(seed=5, n=32)
    assert len(beams) == 5 and all(abs(b[2] - 0.1) < 1e-15 for b in beams)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_fisher_inverse_on_the_host():
    """rte_fisher_inverse (host FP64, no device): F^-1 from the Cholesky factor
    of a random SPD matrix equals numpy's inverse, and a non-positive pivot is
    refused."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((18, 18))
    F = A @ A.T + 18 * np.eye(18)
    Cf = np.linalg.cholesky(F)
    Fi = rte.fisher_inverse(Cf)
    assert np.max(np.abs(Fi - np.linalg.inv(F))) < 1e-12 * np.max(np.abs(np.linalg.inv(F)))
    assert np.array_equal(Fi, Fi.T)
    bad = Cf.copy()
    bad[4, 4] = 0.0
    try:
        rte.fisher_inverse(bad)
        assert False, "zero pivot accepted"
    except rte.RteError as e:
        assert e.code == rte.E_ARG
