"""ctypes binding of the FP64 oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP product path.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module; the product package never does.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

STOP_LAMBDA, STOP_DELTA = 0, 1
QUAD_REF, QUAD_EXACT = 0, 1


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double),
                ("ymin", C.c_double), ("ymax", C.c_double),
                ("zmin", C.c_double), ("zmax", C.c_double)]


class Optics(C.Structure):
    _fields_ = [("mu_a", C.POINTER(C.c_double)), ("mu_s", C.POINTER(C.c_double)),
                ("g", C.c_double), ("n", C.c_double), ("c", C.c_double)]


class Beam(C.Structure):
    _fields_ = [("pos_x", C.c_double), ("pos_y", C.c_double),
                ("diam", C.c_double), ("mag", C.c_double)]


class Params(C.Structure):
    _fields_ = [("L", C.c_int), ("max_terms", C.c_int), ("stop_rule", C.c_int),
                ("tol", C.c_double), ("prop_thresh", C.c_double), ("exit_quad", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("terms", C.c_int), ("lambda_last", C.c_double),
                ("rel_change_last", C.c_double), ("ballistic_total", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc.so is not built (make -f oracle/Makefile)")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.orc_ylm_complex.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp]
        L.orc_ylm_real.argtypes = [C.c_int, C.c_double, C.c_double, dp]
        L.orc_assoc_legendre.argtypes = [C.c_double, C.c_int, C.c_int]
        L.orc_assoc_legendre.restype = C.c_double
        L.orc_self_block.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int, dp]
        L.orc_self_moments.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.orc_self_beta.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.orc_tau.argtypes = [C.POINTER(Grid), dp, C.c_int, C.c_int]
        L.orc_tau.restype = C.c_double
        L.orc_source.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam), C.c_int, dp, dp]
        L.orc_propagate.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_double, dp, dp]
        L.orc_propagate_range.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_double, dp, dp,
                                          C.c_int, C.c_int]
        L.orc_scatter.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.c_int, dp]
        L.orc_exitance.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_int, dp, dp]
        L.orc_quad_weights.argtypes = [C.c_int, C.c_int, dp]
        L.orc_solve.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam),
                                C.POINTER(Params), dp, C.POINTER(Stats)]
        L.orc_solve_field.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam),
                                      C.POINTER(Params), dp, dp, C.POINTER(Stats)]
        L.orc_solve_images.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam), C.POINTER(Params),
                                       dp, dp, C.POINTER(Stats)]
        L.orc_reflectance.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_int, dp, dp]
        L.orc_quad_weights_refl.argtypes = [C.c_int, C.c_int, dp]
        L.orc_ref_derivative.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam), C.POINTER(Params),
                                         C.POINTER(C.c_int), C.c_int, dp, C.POINTER(Stats)]
        L.orc_neumann.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Params), dp, dp, dp, dp,
                                  C.POINTER(Stats)]
        L.orc_bottom_sum.argtypes = [C.POINTER(Grid), C.c_int, dp]
        L.orc_bottom_sum.restype = C.c_double
        L.orc_self_block_c.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_int, dp]
        L.orc_propagate_c.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_double, C.c_double, dp, dp]
        L.orc_propagate_c_range.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_double, C.c_double, dp,
                                            dp, C.c_int, C.c_int]
        L.orc_source_c.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam), C.c_int, C.c_double, dp, dp]
        L.orc_bottom_sum_c.argtypes = [C.POINTER(Grid), C.c_int, dp]
        L.orc_bottom_sum_c.restype = C.c_double
        L.orc_solve_fd.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam), C.POINTER(Params),
                                   C.c_double, dp, dp, dp, C.POINTER(Stats)]
        L.orc_propagate_range_timed.argtypes = [C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_double, dp, dp,
                                                C.c_int, C.c_int, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        _LIB = L
    return _LIB


def set_threads(n):
    """Threads of the oracle's OpenMP loops (n > 0)."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class Problem:
    """One medium on one grid, in the oracle's conventions (x fastest)."""

    def __init__(self, shape, bounds, mu_a, mu_s, g, n=1.0, c=2.9979e10):
        nx, ny, nz = shape
        self.grid = Grid(nx, ny, nz, *bounds)
        N = nx * ny * nz
        self.mu_a = _f64(np.broadcast_to(mu_a, (N,)) if np.ndim(mu_a) == 0 else np.ravel(mu_a))
        self.mu_s = _f64(np.broadcast_to(mu_s, (N,)) if np.ndim(mu_s) == 0 else np.ravel(mu_s))
        assert self.mu_a.size == N and self.mu_s.size == N
        self.optics = Optics(_p(self.mu_a), _p(self.mu_s), g, n, c)
        self.shape = (nx, ny, nz)
        self.N = N

    def solve(self, beam, L, stop_rule=STOP_LAMBDA, tol=1e-6, max_terms=1000,
              prop_thresh=1e30, exit_quad=QUAD_REF, field=False, reflectance=False):
        nx, ny, _ = self.shape
        img = np.zeros(nx * ny)
        st = Stats()
        prm = Params(L, max_terms, stop_rule, tol, prop_thresh, exit_quad)
        b = Beam(*beam)
        refl = None
        if reflectance:
            refl = np.zeros(nx * ny)
            rc = lib().orc_solve_images(C.byref(self.grid), C.byref(self.optics), C.byref(b), C.byref(prm),
                                        _p(img), _p(refl), C.byref(st))
        elif field:
            w = np.zeros((L + 1) ** 2 * self.N)
            rc = lib().orc_solve_field(C.byref(self.grid), C.byref(self.optics), C.byref(b),
                                       C.byref(prm), _p(img), _p(w), C.byref(st))
        else:
            rc = lib().orc_solve(C.byref(self.grid), C.byref(self.optics), C.byref(b),
                                 C.byref(prm), _p(img), C.byref(st))
        if rc != 0:
            raise RuntimeError("orc_solve failed: %d" % rc)
        out = dict(exitance=img.reshape(ny, nx), terms=st.terms, lam=st.lambda_last,
                   rel=st.rel_change_last, ballistic_total=st.ballistic_total)
        if field:
            out["w"] = w.reshape((L + 1) ** 2, self.N)
        if reflectance:
            out["reflectance"] = refl.reshape(ny, nx)
        return out

    def ref_derivative(self, beam, L, region, which, stop_rule=STOP_LAMBDA, tol=1e-6, max_terms=1000,
                       prop_thresh=1e30, exit_quad=QUAD_REF):
        """The reference's forward-mode derivative image of the exitance with
        respect to mu_a (which=0) or mu_s (which=1) on the voxels where region != 0."""
        nx, ny, _ = self.shape
        reg = np.ascontiguousarray(np.ravel(region), dtype=np.int32)
        assert reg.size == self.N
        img = np.zeros(nx * ny)
        st = Stats()
        prm = Params(L, max_terms, stop_rule, tol, prop_thresh, exit_quad)
        rc = lib().orc_ref_derivative(C.byref(self.grid), C.byref(self.optics), C.byref(Beam(*beam)), C.byref(prm),
                                      reg.ctypes.data_as(C.POINTER(C.c_int)), which, _p(img), C.byref(st))
        if rc != 0:
            raise RuntimeError("orc_ref_derivative failed: %d" % rc)
        return img.reshape(ny, nx), st.terms

    def neumann(self, src, L, stop_rule=STOP_LAMBDA, tol=1e-6, max_terms=1000, prop_thresh=1e30,
                exit_quad=QUAD_REF):
        """Neumann(src, out, 1) on a given first term src [K, N]: (field A sum_q W_q
        [K, N], exitance, reflectance, terms), the images without a ballistic term."""
        nx, ny, _ = self.shape
        K = (L + 1) ** 2
        src = _f64(np.reshape(src, (K * self.N,)))
        w = np.zeros(K * self.N)
        img = np.zeros(nx * ny)
        refl = np.zeros(nx * ny)
        st = Stats()
        prm = Params(L, max_terms, stop_rule, tol, prop_thresh, exit_quad)
        rc = lib().orc_neumann(C.byref(self.grid), C.byref(self.optics), C.byref(prm), _p(src), _p(w), _p(img),
                               _p(refl), C.byref(st))
        if rc != 0:
            raise RuntimeError("orc_neumann failed: %d" % rc)
        return w.reshape(K, self.N), img.reshape(ny, nx), refl.reshape(ny, nx), st.terms

    def source(self, beam, L):
        K = (L + 1) ** 2
        s0 = np.zeros(K * self.N)
        ball = np.zeros(self.shape[0] * self.shape[1])
        cols = lib().orc_source(C.byref(self.grid), C.byref(self.optics), C.byref(Beam(*beam)),
                                L, _p(s0), _p(ball))
        return s0.reshape(K, self.N), ball, cols

    def propagate(self, W, L, prop_thresh=1e30):
        W = _f64(W)
        out = np.zeros_like(W)
        lib().orc_propagate(C.byref(self.grid), C.byref(self.optics), L, prop_thresh, _p(W), _p(out))
        return out

    def propagate_range(self, W, L, d_begin, d_end, prop_thresh=1e30):
        W = _f64(W)
        out = np.zeros_like(W)
        lib().orc_propagate_range(C.byref(self.grid), C.byref(self.optics), L, prop_thresh, _p(W), _p(out),
                                  d_begin, d_end)
        return out

    def propagate_range_timed(self, W, L, d_begin, d_end, prop_thresh=1e30):
        """(out, seconds of the pair loop alone) of orc_propagate_range."""
        W = _f64(W)
        out = np.zeros_like(W)
        t = C.c_double(0.0)
        lib().orc_propagate_range_timed(C.byref(self.grid), C.byref(self.optics), L, prop_thresh, _p(W), _p(out),
                                        d_begin, d_end, C.byref(t))
        return out, t.value

    def scatter(self, W, L):
        W = _f64(W).copy()
        lib().orc_scatter(C.byref(self.grid), C.byref(self.optics), L, _p(W))
        return W

    def reflectance(self, w, L, quad=QUAD_REF):
        img = np.zeros(self.shape[0] * self.shape[1])
        lib().orc_reflectance(C.byref(self.grid), C.byref(self.optics), L, quad, _p(_f64(w)), _p(img))
        return img

    def exitance(self, w, L, quad=QUAD_REF):
        img = np.zeros(self.shape[0] * self.shape[1])
        lib().orc_exitance(C.byref(self.grid), C.byref(self.optics), L, quad, _p(_f64(w)), _p(img))
        return img

    def tau(self, d, s):
        mut = _f64(self.mu_a + self.mu_s)
        return lib().orc_tau(C.byref(self.grid), _p(mut), d, s)

    def bottom_sum(self, W, L):
        return lib().orc_bottom_sum(C.byref(self.grid), L, _p(_f64(W)))

    # -- frequency domain: complex mu_t = mu_a + mu_s + i kim, kim = omega n / c --
    def kim(self, omega):
        return omega * self.optics.n / self.optics.c

    def solve_fd(self, beam, L, omega, stop_rule=STOP_LAMBDA, tol=1e-6, max_terms=1000, prop_thresh=1e30,
                 exit_quad=QUAD_REF, field=False):
        """Complex exitance and reflectance of an intensity-modulated beam at
        angular frequency omega (rad/s)."""
        nx, ny, _ = self.shape
        K = (L + 1) ** 2
        img = np.zeros(nx * ny, dtype=np.complex128)
        refl = np.zeros(nx * ny, dtype=np.complex128)
        w = np.zeros(K * self.N, dtype=np.complex128) if field else None
        st = Stats()
        prm = Params(L, max_terms, stop_rule, tol, prop_thresh, exit_quad)
        rc = lib().orc_solve_fd(C.byref(self.grid), C.byref(self.optics), C.byref(Beam(*beam)), C.byref(prm),
                                self.kim(omega), _p(img), _p(refl), _p(w) if field else None, C.byref(st))
        if rc != 0:
            raise RuntimeError("orc_solve_fd failed: %d" % rc)
        out = dict(exitance=img.reshape(ny, nx), reflectance=refl.reshape(ny, nx), terms=st.terms,
                   lam=st.lambda_last, rel=st.rel_change_last, ballistic_total=st.ballistic_total)
        if field:
            out["w"] = w.reshape(K, self.N)
        return out

    def source_c(self, beam, L, omega):
        K = (L + 1) ** 2
        s0 = np.zeros(K * self.N, dtype=np.complex128)
        ball = np.zeros(self.shape[0] * self.shape[1], dtype=np.complex128)
        cols = lib().orc_source_c(C.byref(self.grid), C.byref(self.optics), C.byref(Beam(*beam)), L,
                                  self.kim(omega), _p(s0), _p(ball))
        return s0.reshape(K, self.N), ball, cols

    def propagate_c(self, W, L, omega, prop_thresh=1e30):
        W = _c128(W)
        out = np.zeros_like(W)
        lib().orc_propagate_c(C.byref(self.grid), C.byref(self.optics), L, prop_thresh, self.kim(omega), _p(W),
                              _p(out))
        return out

    def propagate_c_range(self, W, L, omega, d_begin, d_end, prop_thresh=1e30):
        W = _c128(W)
        out = np.zeros_like(W)
        lib().orc_propagate_c_range(C.byref(self.grid), C.byref(self.optics), L, prop_thresh, self.kim(omega), _p(W),
                                    _p(out), d_begin, d_end)
        return out

    def bottom_sum_c(self, W, L):
        return lib().orc_bottom_sum_c(C.byref(self.grid), L, _p(_c128(W)))


def quad_weights_refl(L, quad=QUAD_REF):
    R = np.zeros(L + 1)
    lib().orc_quad_weights_refl(L, quad, _p(R))
    return R


def ylm_complex(L, theta, phi):
    K = (L + 1) ** 2
    re, im = np.zeros(K), np.zeros(K)
    lib().orc_ylm_complex(L, theta, phi, _p(re), _p(im))
    return re + 1j * im


def ylm_real(L, theta, phi):
    out = np.zeros((L + 1) ** 2)
    lib().orc_ylm_real(L, theta, phi, _p(out))
    return out


def self_block(L, dx, dy, dz, mu_t, c_m=1.0, method=0):
    K = (L + 1) ** 2
    S = np.zeros(K * K)
    lib().orc_self_block(L, dx, dy, dz, mu_t, c_m, method, _p(S))
    return S.reshape(K, K)


def self_block_c(L, dx, dy, dz, mu_t, c_m=1.0, method=0):
    """S(mu_t) for complex mu_t (interleaved in C, complex128 here)."""
    K = (L + 1) ** 2
    S = np.zeros(K * K, dtype=np.complex128)
    lib().orc_self_block_c(L, dx, dy, dz, mu_t.real, mu_t.imag, c_m, method, _p(S))
    return S.reshape(K, K)


def self_beta(dx, dy, dz):
    b = np.zeros(40000)
    lib().orc_self_beta(dx, dy, dz, _p(b))
    return b


def quad_weights(L, quad=QUAD_REF):
    Q = np.zeros(L + 1)
    lib().orc_quad_weights(L, quad, _p(Q))
    return Q
