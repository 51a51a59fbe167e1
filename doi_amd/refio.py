"""The reference's file formats, so its parameter files run unchanged.

* Parameter files: INI text read with the rules of the iniparser library the
  reference links (CPU package/src/iniparser.c:455-511): `[section]` lines,
  `key = value` lines (value in double or single quotes, or up to a `;`/`#`
  comment), `#`/`;` comment lines; section and key names are case-insensitive
  and looked up as "section:key"; a key given twice keeps the last value.
  Numbers are read like atof / strtol(base 0): the longest numeric prefix.
* Problem set-up: keys of CPU package/src/Inits.c:16-134 and
  CPU package/src/RTE_uniform.c:45-60 (uniform mua/mus), or of
  GPU package/src/inits.cu:22-123 and rte.cu:100-120 (NoTiss, TissAbs<i>,
  TissSc<i>, TissMapFile: one byte per voxel, x fastest, tissue 0 = outside).
* Images: OutputImage (CPU package/src/Dist.c:16-40): int32 nX, int32 nY, then
  nX*nY doubles of the real part and nX*nY of the imaginary part, x fastest.
"""
import os
import re
import struct

import numpy as np

_NUM = re.compile(r"^[ \t]*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)")
_INT = re.compile(r"^[ \t]*[+-]?(0[xX][0-9a-fA-F]+|0[0-7]*|[1-9]\d*)")


def parse_ini(text):
    """{'section:key': 'value'} (names lowercased; sections also as 'section')."""
    d = {}
    section = ""
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        line = lines[i].strip()
        # continuation lines end with a backslash
        while line.endswith("\\") and i + 1 < len(lines):
            i += 1
            line = line[:-1].rstrip() + lines[i].strip()
        i += 1
        if not line or line[0] in "#;":
            continue
        if line[0] == "[":
            end = line.find("]")
            section = (line[1:end] if end > 0 else line[1:]).strip().lower()
            d[section] = None
            continue
        if "=" not in line:
            continue
        key, val = line.split("=", 1)
        key = key.strip().lower()
        val = val.strip()
        if len(val) >= 2 and val[0] in "\"'" and val.find(val[0], 1) > 0:
            val = val[1:val.find(val[0], 1)]
        else:
            cut = min([p for p in (val.find(";"), val.find("#")) if p >= 0], default=len(val))
            val = val[:cut].strip()
        d[(section + ":" + key) if section else key] = val
    return d


def read_ini(path):
    with open(path, "r", errors="replace") as f:
        return parse_ini(f.read())


def get_double(ini, key, default):
    v = ini.get(key.lower())
    if v is None:
        return default
    m = _NUM.match(v)
    return float(m.group(0)) if m else 0.0


def get_int(ini, key, default):
    v = ini.get(key.lower())
    if v is None:
        return default
    m = _INT.match(v)
    return int(m.group(0).strip(), 0) if m else 0


def get_string(ini, key, default=None):
    v = ini.get(key.lower())
    return default if v is None else v


def read_tissue_map(path, shape):
    """Byte tissue map, x fastest (GPU package/src/inits.cu:97-123)."""
    nx, ny, nz = shape
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < nx * ny * nz:
        raise ValueError("tissue map %s holds %d voxels, the grid needs %d" % (path, raw.size, nx * ny * nz))
    return raw[:nx * ny * nz].copy()


def write_image(path, real, imag=None):
    """OutputImage: int32 nX, int32 nY, real plane, imaginary plane (x fastest)."""
    real = np.asarray(real, dtype=np.float64)
    ny, nx = real.shape
    im = np.zeros_like(real) if imag is None else np.asarray(imag, dtype=np.float64).reshape(ny, nx)
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", nx, ny))
        f.write(np.ascontiguousarray(real).astype("<f8").tobytes())
        f.write(np.ascontiguousarray(im).astype("<f8").tobytes())


def write_phase_image(path, phase):
    """OutputPhaseImage: int32 nX, int32 nY, one plane of doubles, x fastest
    (CPU package/src/Dist.c:46-65)."""
    phase = np.asarray(phase, dtype=np.float64)
    ny, nx = phase.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", nx, ny))
        f.write(np.ascontiguousarray(phase).astype("<f8").tobytes())


def read_phase_image(path):
    raw = open(path, "rb").read()
    nx, ny = struct.unpack("<ii", raw[:8])
    return np.frombuffer(raw[8:8 + 8 * nx * ny], dtype="<f8").reshape(ny, nx).copy()


def read_image(path):
    """(real, imag) planes of an OutputImage file, each [ny][nx]."""
    raw = open(path, "rb").read()
    nx, ny = struct.unpack("<ii", raw[:8])
    v = np.frombuffer(raw[8:8 + 16 * nx * ny], dtype="<f8")
    return v[:nx * ny].reshape(ny, nx).copy(), v[nx * ny:].reshape(ny, nx).copy()


class Problem:
    """What a reference parameter file describes: grid, optical maps, beam and
    algorithm settings, plus the output file names it asks for."""

    def __init__(self, ini, base_dir="."):
        self.ini = ini
        g = ini
        self.shape = (get_int(g, "Geometry:nX", -1), get_int(g, "Geometry:nY", -1), get_int(g, "Geometry:nZ", -1))
        if min(self.shape) < 1:
            raise ValueError("[Geometry] nX, nY and nZ must be given and >= 1")
        self.bounds = tuple(get_double(g, "Geometry:" + k, -1.0)
                            for k in ("xMin", "xMax", "yMin", "yMax", "zMin", "zMax"))
        self.prop_thresh = get_double(g, "Geometry:propThresh", 0.0)
        self.g = get_double(g, "Phantom:g", -1.0)
        self.n = get_double(g, "Phantom:n", -1.0)
        self.c = get_double(g, "Phantom:c", 2.9979e10)
        N = self.shape[0] * self.shape[1] * self.shape[2]
        map_file = get_string(g, "Phantom:TissMapFile")
        if map_file is not None or get_string(g, "Phantom:NoTiss") is not None:
            # tissue map + per-tissue coefficients (GPU package)
            ntiss = get_int(g, "Phantom:NoTiss", 1) + 1
            mua = np.zeros(ntiss)
            mus = np.zeros(ntiss)
            for t in range(1, ntiss):
                mua[t] = get_double(g, "Phantom:TissAbs%d" % t, 0.01)
                mus[t] = get_double(g, "Phantom:TissSc%d" % t, 1.0)
            if map_file is None:
                tiss = np.ones(N, dtype=np.uint8)
            else:
                path = map_file if os.path.isabs(map_file) or os.path.exists(map_file) else \
                    os.path.join(base_dir, map_file)
                tiss = read_tissue_map(path, self.shape)
            if np.any(tiss >= ntiss):
                raise ValueError("tissue map holds type %d, NoTiss is %d" % (int(tiss.max()), ntiss - 1))
            if np.any(tiss == 0):
                raise ValueError("tissue map holds type 0 (outside, mu = 0) inside the grid")
            self.tissue = tiss
            self.mu_a = mua[tiss]
            self.mu_s = mus[tiss]
        else:
            self.tissue = None
            self.mu_a = np.full(N, get_double(g, "Phantom:mua", -1.0))
            self.mu_s = np.full(N, get_double(g, "Phantom:mus", -1.0))
        # the beam keeps the reference's +1e-16 on the diameter inside the library
        self.beam = (get_double(g, "Source:posX", 0.0), get_double(g, "Source:posY", 0.0),
                     get_double(g, "Source:diam", 0.2), get_double(g, "Source:magnitude", 1.0))
        self.modulation = get_double(g, "Source:modulation", 0.0)
        self.L = get_int(g, "Algorithm:nL", -1)
        if self.L < 0:
            raise ValueError("[Algorithm] nL must be given")
        self.n_terms = get_int(g, "Algorithm:nTerms", 1)
        self.trans_file = get_string(g, "Runtime:TransFile")
        self.reflect_file = get_string(g, "Runtime:ReflectFile")
        self.phase_file = get_string(g, "Runtime:PhaseFile")


def load(path):
    return Problem(read_ini(path), base_dir=os.path.dirname(os.path.abspath(path)))
