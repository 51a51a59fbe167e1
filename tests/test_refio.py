"""Reference file formats (doi_amd.refio): the INI rules of the iniparser the
reference links, the tissue-map and OutputImage formats, and parameter files
written the way the reference CPU package was run for the golden fixtures:
loaded into a problem and solved by the oracle they give the reference
binary's images (CPU only)."""
import struct

import numpy as np
import pytest

from doi_amd import refio
from oracle import oracle as O
from tests.helpers import rel_max

INI = """[Runtime]
TransFile = trans.out
ReflectFile = refl.out
PhaseFile = phase.out

[Algorithm]
nL = {L}
nTerms = 30

[Geometry]
nX = {nx}
nY = {ny}
nZ = {nz}
xMin = {b[0]!r}
xMax = {b[1]!r}
yMin = {b[2]!r}
yMax = {b[3]!r}
zMin = {b[4]!r}
zMax = {b[5]!r}
propThresh = {pt!r}
selfSubVox = 1

[Phantom]
mua = {mua!r}
mus = {mus!r}
g = {g!r}
n = {n!r}

[Source]
posX = {beam[0]!r}
posY = {beam[1]!r}
diam = {beam[2]!r}
magnitude = 1.0
modulation = {mod!r}
"""


def ini_for(case, mod=0.0):
    nx, ny, nz = case["shape"]
    b = case["beam"]
    return INI.format(L=case["L"], nx=nx, ny=ny, nz=nz, b=case["bounds"], pt=case["prop_thresh"], mua=case["mua"],
                      mus=case["mus"], g=case["g"], n=case["n"], beam=(b["pos_x"], b["pos_y"], b["diam"]), mod=mod)


def test_ini_rules():
    d = refio.parse_ini("""
# comment
; another
[Geometry]
NX = 8 ; inline comment
nY=0x10
nZ = 7.9
xMin = -0.4abc
Name = "quoted ; not a comment"
Other = 'single'
long = 1, \\
  2
[phantom]
mua = 0.01
mua = 0.02
""")
    assert refio.get_int(d, "Geometry:nX", -1) == 8
    assert refio.get_int(d, "geometry:NY", -1) == 16
    assert refio.get_int(d, "Geometry:nZ", -1) == 7
    assert refio.get_double(d, "Geometry:xMin", 0.0) == -0.4
    assert refio.get_string(d, "Geometry:name") == "quoted ; not a comment"
    assert refio.get_string(d, "Geometry:other") == "single"
    assert refio.get_string(d, "Geometry:long") == "1,2"
    assert refio.get_double(d, "Phantom:mua", 0.0) == 0.02
    assert refio.get_double(d, "Phantom:missing", 5.5) == 5.5


def test_image_round_trip_and_layout(tmp_path):
    real = np.arange(12, dtype=np.float64).reshape(3, 4) * 0.5
    p = tmp_path / "img.out"
    refio.write_image(str(p), real)
    raw = p.read_bytes()
    assert len(raw) == 8 + 2 * 12 * 8
    assert struct.unpack("<ii", raw[:8]) == (4, 3)
    assert struct.unpack("<d", raw[8 + 8:8 + 16])[0] == 0.5   # x fastest
    re_, im = refio.read_image(str(p))
    assert np.array_equal(re_, real) and not np.any(im)


def test_tissue_map_problem(tmp_path):
    shape = (4, 3, 5)
    N = 60
    tiss = (np.arange(N) % 3 + 1).astype(np.uint8)
    tiss.tofile(str(tmp_path / "map.bin"))
    text = """[Algorithm]
nL = 2
[Geometry]
nX = 4
nY = 3
nZ = 5
xMin = -0.2
xMax = 0.2
yMin = -0.15
yMax = 0.15
zMin = 0
zMax = 0.5
[Phantom]
g = 0.8
n = 1.0
NoTiss = 3
TissAbs1 = 0.01
TissSc1 = 1.0
TissAbs2 = 0.02
TissSc2 = 1.5
TissAbs3 = 0.05
TissMapFile = map.bin
[Source]
posX = 0.05
"""
    (tmp_path / "p.ini").write_text(text)
    prob = refio.load(str(tmp_path / "p.ini"))
    assert prob.shape == shape and prob.L == 2 and prob.bounds == (-0.2, 0.2, -0.15, 0.15, 0.0, 0.5)
    mua = np.array([0.0, 0.01, 0.02, 0.05])[tiss]
    mus = np.array([0.0, 1.0, 1.5, 1.0])[tiss]          # TissSc3 defaults to 1.0
    assert np.array_equal(prob.mu_a, mua) and np.array_equal(prob.mu_s, mus)
    assert prob.beam == (0.05, 0.0, 0.2, 1.0)             # reference defaults
    bad = tiss.copy()
    bad[7] = 0
    bad.tofile(str(tmp_path / "map.bin"))
    with pytest.raises(ValueError):
        refio.load(str(tmp_path / "p.ini"))


def test_parameter_files_reproduce_reference_images(tmp_path, golden_ref):
    """The golden cases' parameter files through refio and the oracle give the
    reference binary's images and term counts."""
    for case in golden_ref[:4] + golden_ref[6:]:
        p = tmp_path / (case["name"] + ".ini")
        p.write_text(ini_for(case))
        prob = refio.load(str(p))
        P = O.Problem(prob.shape, prob.bounds, prob.mu_a, prob.mu_s, prob.g, n=prob.n, c=prob.c)
        r = P.solve(prob.beam, prob.L, stop_rule=O.STOP_LAMBDA, max_terms=1000, prop_thresh=prob.prop_thresh)
        assert r["terms"] == case["terms"], case["name"]
        assert rel_max(r["exitance"], np.array(case["trans_image"])) < 1e-12, case["name"]


def test_face_quadratures_reproduce_reference_photon_totals(golden_ref):
    """doi_amd.faces on the oracle's final field reproduces the photon total
    the reference binary printed (%f, 6 decimals) for every golden case; its
    real harmonics equal the oracle's."""
    from doi_amd import faces
    rng = np.random.default_rng(1)
    th, ph = rng.uniform(0, np.pi, 6), rng.uniform(0, 2 * np.pi, 6)
    R = faces.real_sh(6, th, ph)
    for k in range(6):
        assert np.max(np.abs(R[:, k] - O.ylm_real(6, th[k], ph[k]))) < 1e-13
    for c in golden_ref:
        b = c["beam"]
        P = O.Problem(tuple(c["shape"]), tuple(c["bounds"]), c["mua"], c["mus"], c["g"], n=c["n"])
        r = P.solve((b["pos_x"], b["pos_y"], b["diam"], b["mag"]), c["L"], prop_thresh=c["prop_thresh"],
                    field=True)
        tot = faces.photon_number(r["w"], tuple(c["shape"]), tuple(c["bounds"]), c["L"], r["ballistic_total"],
                                  n=c["n"])
        assert abs(tot - c["photons_printed"]) <= 5.0e-7 + 1e-12, c["name"]


def test_side_face_images_reproduce_reference_files(golden_ref, tmp_path):
    """doi_amd.faces.side_face_images on the oracle's final field against the
    SideFace1..4 files the reference binary wrote (cubic grids, where its header
    is the face's shape), and the files written back in that layout."""
    from doi_amd import faces
    n_checked = 0
    for c in golden_ref:
        if "side_faces" not in c:
            continue
        b = c["beam"]
        P = O.Problem(tuple(c["shape"]), tuple(c["bounds"]), c["mua"], c["mus"], c["g"], n=c["n"])
        r = P.solve((b["pos_x"], b["pos_y"], b["diam"], b["mag"]), c["L"], prop_thresh=c["prop_thresh"],
                    field=True)
        imgs = faces.side_face_images(r["w"], tuple(c["shape"]), tuple(c["bounds"]), c["L"], n=c["n"])
        for face, _ in faces.SIDE_FACE_FILES:
            ref = np.array(c["side_faces"][face])
            assert c["side_faces"][face + "_imag_max"] < 1e-15
            assert imgs[face].shape == ref.shape
            assert rel_max(imgs[face], ref) < 1e-10, (c["name"], face)
        n_checked += 1
        if n_checked == 1:
            faces.write_side_faces(str(tmp_path), imgs)
            for face, name in faces.SIDE_FACE_FILES:
                re_, im = refio.read_image(str(tmp_path / name))
                assert np.array_equal(re_, imgs[face]) and not np.any(im)
    assert n_checked == 6
