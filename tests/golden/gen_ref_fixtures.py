"""Writes tests/golden/ref_cpu_uniform.json (exitance images and Neumann term
counts produced by the reference CPU package, uniform media, FP64) and
tests/golden/ref_cpu_modulated.json (the same binary with a modulated source:
transmitted, reflected and phase images, every printed term count), built
from its own sources by oracle/Makefile.ref into oracle/_ref/rte_cpu.

Run in the build container only (it needs /root/reference to build the
binary):  make -f oracle/Makefile.ref && python tests/golden/gen_ref_fixtures.py
The fixture holds data only: the .par inputs and the binary's outputs
(transmitted and reflected images, term count, printed photon total, and for
cubic grids the four side-face images).
"""
import json
import os
import re
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(REPO, "oracle", "_ref", "rte_cpu")

# name, grid (nx,ny,nz), bounds, mua, mus, g, n, L, beam (posX,posY,diam), propThresh
CASES = [
    ("cube8_L3_g0.8", (8, 8, 8), (-0.4, 0.4, -0.4, 0.4, 0.0, 0.8), 0.01, 1.0, 0.8, 1.0, 3, (0.0, 0.0, 0.2), 10.0),
    ("cube8_L0_g0.8", (8, 8, 8), (-0.4, 0.4, -0.4, 0.4, 0.0, 0.8), 0.01, 1.0, 0.8, 1.0, 0, (0.0, 0.0, 0.2), 10.0),
    ("cube8_L1_g0.8", (8, 8, 8), (-0.4, 0.4, -0.4, 0.4, 0.0, 0.8), 0.01, 1.0, 0.8, 1.0, 1, (0.0, 0.0, 0.2), 10.0),
    ("cube8_L5_g0.8", (8, 8, 8), (-0.4, 0.4, -0.4, 0.4, 0.0, 0.8), 0.01, 1.0, 0.8, 1.0, 5, (0.0, 0.0, 0.2), 10.0),
    ("cube8_L3_g0", (8, 8, 8), (-0.4, 0.4, -0.4, 0.4, 0.0, 0.8), 0.01, 1.0, 0.0, 1.0, 3, (0.0, 0.0, 0.2), 10.0),
    ("cube8_L1_g0", (8, 8, 8), (-0.4, 0.4, -0.4, 0.4, 0.0, 0.8), 0.01, 1.0, 0.0, 1.0, 1, (0.0, 0.0, 0.2), 10.0),
    ("box_6x5x4_L2_offcentre_n1.37", (6, 5, 4), (-0.3, 0.3, -0.2, 0.3, 0.0, 0.48), 0.05, 2.0, 0.5, 1.37, 2,
     (0.1, 0.05, 0.15), 10.0),
    ("slab_7x7x3_L4_g0.9_wide_beam", (7, 7, 3), (-0.35, 0.35, -0.35, 0.35, 0.0, 0.3), 0.1, 3.0, 0.9, 1.0, 4,
     (0.0, 0.0, 0.35), 10.0),
]

# modulated sources (CPU package/src/RTE_uniform.c:120-177): name, case of CASES it
# modifies, modulation (angular frequency, rad/s)
MOD_CASES = [
    ("cube8_L3_g0.8_mod6.3e9", "cube8_L3_g0.8", 6.283185307179586e9),
    ("cube8_L1_g0_mod3e10", "cube8_L1_g0", 3.0e10),
    ("box_6x5x4_L2_offcentre_n1.37_mod2e10", "box_6x5x4_L2_offcentre_n1.37", 2.0e10),
]

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
subVox = 0
subThresh = 0.0
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


def run_case(case, mod=0.0):
    name, (nx, ny, nz), b, mua, mus, g, n, L, beam, pt = case
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "in.ini"), "w") as f:
            f.write(INI.format(L=L, nx=nx, ny=ny, nz=nz, b=b, pt=pt, mua=mua, mus=mus, g=g, n=n, beam=beam,
                               mod=mod))
        out = subprocess.run([BIN, "in.ini"], cwd=d, capture_output=True, text=True, check=True, timeout=600,
                             env=dict(os.environ, OMP_NUM_THREADS="8")).stdout
        counts = [int(t) for t in re.findall(r"Neumann series: (\d+)", out)]
        photons = float(re.findall(r"Number of photons total: ([-+0-9.eE]+)", out)[-1])
        images = {}
        for key, fname in (("trans_image", "trans.out"), ("refl_image", "refl.out")):
            raw = open(os.path.join(d, fname), "rb").read()
            rx, ry = struct.unpack("ii", raw[:8])
            vals = np.frombuffer(raw[8:], dtype=np.float64)
            images[key] = vals[: rx * ry].reshape(ry, rx).tolist()
            if mod:
                images[key + "_imag"] = vals[rx * ry: 2 * rx * ry].reshape(ry, rx).tolist()
        if not mod and nx == ny == nz:
            # ComputeSideFaceImages writes these under the OutputImage header (nX, nY):
            # only a cube gives the face's own shape (CPU package/src/Dist.c:409-522)
            side = {}
            for face, fname in (("-x", "SideFace1(-X).out"), ("+x", "SideFace2(+X).out"),
                                ("-y", "SideFace3(-Y).out"), ("+y", "SideFace4(+Y).out")):
                raw = open(os.path.join(d, fname), "rb").read()
                rx, ry = struct.unpack("ii", raw[:8])
                vals = np.frombuffer(raw[8:], dtype=np.float64)
                side[face] = vals[: rx * ry].reshape(ry, rx).tolist()
                side[face + "_imag_max"] = float(np.max(np.abs(vals[rx * ry: 2 * rx * ry])))
            images["side_faces"] = side
        if mod:
            raw = open(os.path.join(d, "phase.out"), "rb").read()
            rx, ry = struct.unpack("ii", raw[:8])
            images["phase_image"] = np.frombuffer(raw[8:], dtype=np.float64)[: rx * ry].reshape(ry, rx).tolist()
    res = dict(name=name, shape=[nx, ny, nz], bounds=list(b), mua=mua, mus=mus, g=g, n=n, L=L,
               beam=dict(pos_x=beam[0], pos_y=beam[1], diam=beam[2], mag=1.0), prop_thresh=pt,
               terms=counts[-1], photons_printed=photons, **images)
    if mod:
        res.update(modulation=mod, terms_printed=counts)
    return res


def main():
    res = [run_case(c) for c in CASES]
    byname = {c[0]: c for c in CASES}
    mod = []
    for name, base, om in MOD_CASES:
        r = run_case((name,) + tuple(byname[base][1:]), mod=om)
        mod.append(r)
        print(r["name"], r["terms_printed"])
    with open(os.path.join(HERE, "ref_cpu_modulated.json"), "w") as f:
        json.dump(dict(source="reference CPU package (CPU package/src), built by oracle/Makefile.ref, "
                              "[Source] modulation != 0",
                       image_layout="[j][i] (OutputImage order); trans/refl _imag: the imaginary planes; "
                                    "phase_image: PhaseFile (degrees); terms_printed: every 'Neumann series' "
                                    "line (CW solve, modulated solve, final)",
                       cases=mod), f, indent=1)
    with open(os.path.join(HERE, "ref_cpu_uniform.json"), "w") as f:
        json.dump(dict(source="reference CPU package (CPU package/src), built by oracle/Makefile.ref",
                       image_layout="trans_image / refl_image [j][i], j = y index, i = x index (OutputImage "
                                    "order); photons_printed: the binary's 'Number of photons total' (%f); "
                                    "side_faces (cubic grids): SideFace1..4 real planes, -x/+x [j][k], "
                                    "-y/+y [i][k] (k = z index fastest), with the largest imaginary value",
                       cases=res), f, indent=1)
    for r in res:
        print(r["name"], r["terms"], np.sum(r["trans_image"]))


if __name__ == "__main__":
    main()
