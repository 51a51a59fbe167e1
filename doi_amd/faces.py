"""Boundary-face fluxes of a final radiance field: the reference's photon
count and side-face images (CPU package/src/Dist.c:349-989), evaluated on
the host from the field rte_solve_field returns (coefficient-major, real SH
basis, the reference's units).

Each face integrates the outgoing radiance of its boundary voxels over the
reference's (theta, phi) grid of num x num points (num = 101 in the
reference's main, CPU package/src/RTE_uniform.c:186), left Riemann sums:

  bottom (z = zmax): theta = i (pi/2)/num, phi = j 2pi/num, weight sin(t) cos(t)
  top    (z = zmin): theta = pi/2 + i (pi/2)/num, same phi, weight -sin(t) cos(t)
  -x / +x faces:     theta = i pi/num, phi = pi/2 + j pi/num / -pi/2 + j pi/num,
                     weight -/+ sin(t)^2 cos(p)
  -y / +y faces:     theta = i pi/num, phi = pi + j pi/num / j pi/num,
                     weight -/+ sin(t)^2 sin(p)

times dtheta dphi, the face area of a voxel and c/n.  The inward flux through
a face uses the opposite face's half-grid.  The photon count the reference
prints is the sum of the outward and inward fluxes plus the ballistic
transmission.  The reference evaluates the complex harmonics;
for a real radiance the sum over the complex basis equals the sum over the
real one at every direction, so the real basis is used here.
"""
import math

import numpy as np

NUM = 101


def real_sh(L, theta, phi):
    """R[a, ...] for a = l^2 + l + m: R_l0 = Y_l0, R_lm = sqrt2 Re Y_lm,
    R_l,-m = sqrt2 Im Y_lm (m > 0), Y_lm orthonormal with the Condon-Shortley
    phase (CPU package/src/SphHarm.c:75-103, 113-160)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    x, s = np.cos(theta), np.sin(theta)
    K = (L + 1) ** 2
    out = np.zeros((K,) + theta.shape)
    for m in range(L + 1):
        # P_m^m = (-1)^m (2m-1)!! s^m, then the upward recurrence in l
        dfact = 1.0
        for i in range(2 * m - 1, 0, -2):
            dfact *= i
        pmm = (-1.0) ** m * dfact * s ** m
        prev, cur = np.zeros_like(x), pmm
        for l in range(m, L + 1):
            if l > m:
                nxt = ((2 * l - 1) * x * cur - (l + m - 1) * prev) / (l - m)
                prev, cur = cur, nxt
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            if m == 0:
                out[l * l + l] = norm * cur
            else:
                out[l * l + l + m] = math.sqrt(2.0) * norm * cur * np.cos(m * phi)
                out[l * l + l - m] = math.sqrt(2.0) * norm * cur * np.sin(m * phi)
    return out


def _face_weights(L, face, num):
    """F[a] = sum over the face's direction grid of R_a times its weight (per unit area and c/n)."""
    i = np.arange(num)[:, None]
    j = np.arange(num)[None, :]
    if face in ("bottom", "top"):
        dth, dph = (math.pi / 2) / num, 2 * math.pi / num
        th = i * (math.pi / 2) / num + (math.pi / 2 if face == "top" else 0.0)
        ph = j * (2 * math.pi) / num
        w = np.sin(th) * np.cos(th) * (1.0 if face == "bottom" else -1.0)
    else:
        dth, dph = math.pi / num, math.pi / num
        th = i * math.pi / num
        off = {"-x": math.pi / 2, "+x": -math.pi / 2, "-y": math.pi, "+y": 0.0}[face]
        ph = j * math.pi / num + off
        trig = np.cos(ph) if face in ("-x", "+x") else np.sin(ph)
        w = np.sin(th) ** 2 * trig * (-1.0 if face in ("-x", "-y") else 1.0)
    th, ph = np.broadcast_arrays(th, ph)
    w = np.broadcast_to(w * dth * dph, th.shape)
    R = real_sh(L, th, ph)
    return np.tensordot(R, w, axes=([1, 2], [0, 1]))


# the face whose outward half-grid measures the inward flux through each face
_OPPOSITE = {"bottom": "top", "top": "bottom", "-x": "+x", "+x": "-x", "-y": "+y", "+y": "-y"}


def face_fluxes(field, shape, bounds, L, c=2.9979e10, n=1.0, num=NUM):
    """(outward, inward) flux images of each face of the grid: dicts face ->
    2-D array (bottom/top [ny][nx], -x/+x [nz][ny], -y/+y [nz][nx]).  The
    inward flux through a face uses the opposite face's half-grid and weight
    (CPU package/src/Dist.c: bottom 564-617 out and 711-731 in, top 650-706 out
    and 622-645 in, -x 736-770 out and 835-851 in, +x 796-830 out and 775-791 in,
    -y 856-890 out and 955-971 in, +y 916-950 out and 895-911 in)."""
    nx, ny, nz = shape
    dx = (bounds[1] - bounds[0]) / nx
    dy = (bounds[3] - bounds[2]) / ny
    dz = (bounds[5] - bounds[4]) / nz
    K = (L + 1) ** 2
    f = np.asarray(field, dtype=np.float64).reshape(K, nz, ny, nx)
    cm = c / n
    W = {face: _face_weights(L, face, num) for face in _OPPOSITE}
    outward, inward = {}, {}
    for face, area, sl in (("bottom", dx * dy, f[:, nz - 1, :, :]), ("top", dx * dy, f[:, 0, :, :]),
                           ("-x", dy * dz, f[:, :, :, 0]), ("+x", dy * dz, f[:, :, :, nx - 1]),
                           ("-y", dx * dz, f[:, :, 0, :]), ("+y", dx * dz, f[:, :, ny - 1, :])):
        outward[face] = np.tensordot(W[face] * area * cm, sl, axes=(0, 0))
        inward[face] = np.tensordot(W[_OPPOSITE[face]] * area * cm, sl, axes=(0, 0))
    return outward, inward


# ComputeSideFaceImages' files (CPU package/src/Dist.c:409, 446, 485, 522)
SIDE_FACE_FILES = (("-x", "SideFace1(-X).out"), ("+x", "SideFace2(+X).out"), ("-y", "SideFace3(-Y).out"),
                   ("+y", "SideFace4(+Y).out"))


def side_face_images(field, shape, bounds, L, c=2.9979e10, n=1.0, num=NUM):
    """The four side-face images as ComputeSideFaceImages lays them out: the
    outward flux of each boundary voxel, -x/+x as [j][k] and -y/+y as [i][k]
    (k, the z index, fastest; CPU package/src/Dist.c:391-406, 428-443, 467-482,
    504-519)."""
    out, _ = face_fluxes(field, shape, bounds, L, c, n, num)
    return {f: np.ascontiguousarray(out[f].T) for f, _ in SIDE_FACE_FILES}


def write_side_faces(directory, images):
    """The reference's four side-face files in `directory`, in the OutputImage
    format with each face's own dimensions as the header (the reference writes
    nX, nY there, the same numbers on a cube)."""
    import os

    from . import refio
    for face, name in SIDE_FACE_FILES:
        refio.write_image(os.path.join(directory, name), images[face])


def photon_number(field, shape, bounds, L, ballistic_total, c=2.9979e10, n=1.0, num=NUM):
    """The reference's photon total: outward plus inward flux through the six
    faces plus the ballistic transmission (CPU package/src/Dist.c:538-988,
    returned as ret + ret_in and printed by CPU package/src/RTE_uniform.c:186)."""
    out, inw = face_fluxes(field, shape, bounds, L, c, n, num)
    return float(ballistic_total + sum(np.sum(v) for v in out.values()) + sum(np.sum(v) for v in inw.values()))
