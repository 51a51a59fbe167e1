"""ctypes binding of librte.so (include/rte.h).

Mirrors the reference solver's entry points: a grid, optical maps, a pencil
beam and the SH order in; the bottom-face exitance out
(CPU package/src/RTE_uniform.c:79-181).  Field arrays at this boundary are
coefficient-major ``W[a, v]`` in the real SH basis, ``v = x + nx*(y + ny*z)``.
"""
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librte.so")

OK, E_ARG, E_DEVICE, E_NOMEM, E_UNSUPP, E_BUSY, E_NOTFOUND = 0, -1, -2, -3, -4, -5, -6
MAX_L = 7
STOP_LAMBDA, STOP_DELTA = 0, 1
PREC_FP64, PREC_MIXED, PREC_FP32 = 0, 1, 2
QUAD_REF, QUAD_EXACT = 0, 1
PATH_CACHED, PATH_MARCH = 0, 1


class RteError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("librte error %d: %s" % (code, msg))
        self.code = code


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
                ("tol", C.c_double), ("prop_thresh", C.c_double), ("precision", C.c_int),
                ("exit_quad", C.c_int), ("path_mode", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("terms", C.c_int), ("lambda_last", C.c_double), ("rel_change_last", C.c_double),
                ("total_exitance", C.c_double), ("ballistic_total", C.c_double),
                ("gpu_ms", C.c_double), ("latency_ms", C.c_double)]


class GradOpts(C.Structure):
    _fields_ = [("checkpoint_stride", C.c_int), ("jacobian", C.c_int)]


class Prof(C.Structure):
    _fields_ = [("prop_launches", C.c_longlong), ("prop_ms", C.c_double), ("prop_pairs", C.c_double),
                ("prop_flops", C.c_double), ("ktab_launches", C.c_longlong), ("ktab_ms", C.c_double),
                ("ktab_pairs", C.c_double), ("other_ms", C.c_double), ("path_mode", C.c_int),
                ("L", C.c_int), ("precision", C.c_int), ("group_launches", C.c_longlong * 6),
                ("group_ms", C.c_double * 6), ("group_flops", C.c_double * 6)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_longlong), ("max_slots_per_bucket", C.c_int), ("max_queue", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_longlong), ("running", C.c_longlong), ("done", C.c_longlong),
                ("failed", C.c_longlong), ("cancelled", C.c_longlong), ("steps", C.c_longlong),
                ("hbm_used", C.c_longlong), ("hbm_budget", C.c_longlong), ("buckets", C.c_int),
                ("worker_status", C.c_int)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = 0, 1, 2, 3, 4


# every entry point of include/rte.h: name -> (restype, argtypes)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
SIGNATURES = {
    "rte_version": (C.c_int, []),
    "rte_last_error": (C.c_char_p, []),
    "rte_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "rte_ctx_destroy": (C.c_int, [_vp]),
    "rte_ctx_profile": (C.c_int, [_vp, C.c_int]),
    "rte_ctx_set_group_size": (C.c_int, [_vp, C.c_int]),
    "rte_ctx_get_profile": (C.c_int, [_vp, C.POINTER(Prof)]),
    "rte_medium_create": (C.c_int, [_vp, C.POINTER(Grid), C.POINTER(Optics), C.POINTER(_vp)]),
    "rte_medium_destroy": (C.c_int, [_vp]),
    "rte_medium_reset_cache": (C.c_int, [_vp]),
    "rte_solve_footprint": (C.c_longlong, [C.POINTER(Grid), C.POINTER(Params), C.c_int]),
    "rte_solve": (C.c_int, [_vp, C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam),
                            C.POINTER(Params), _dp, C.POINTER(Stats)]),
    "rte_solve_media": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(Beam), C.POINTER(Params),
                                  _dp, C.POINTER(Stats)]),
    "rte_solve_images": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(Beam), C.POINTER(Params),
                                   _dp, _dp, C.POINTER(Stats)]),
    "rte_solve_field": (C.c_int, [_vp, _vp, C.POINTER(Beam), C.POINTER(Params), _dp, _dp, _dp, C.POINTER(Stats)]),
    "rte_neumann": (C.c_int, [_vp, _vp, C.POINTER(Params), _dp, _dp, _dp, _dp, C.POINTER(Stats)]),
    "rte_neumann_ref": (C.c_int, [_vp, _vp, C.POINTER(Params), _dp, _dp, C.c_int, C.POINTER(Stats)]),
    "rte_solve_fd": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(Beam), _dp, C.POINTER(Params), _dp, _dp, _dp,
                               C.POINTER(Stats)]),
    "rte_exitance_vjp": (C.c_int, [_vp, _vp, C.POINTER(Beam), C.POINTER(Params), C.POINTER(GradOpts), _dp, _dp,
                                   _dp, _dp, C.POINTER(Stats)]),
    "rte_gradient": (C.c_int, [_vp, _vp, C.POINTER(Beam), C.POINTER(Params), C.POINTER(GradOpts), _dp, _dp, _dp,
                               _dp, _dp, C.POINTER(Stats)]),
    "rte_fisher": (C.c_int, [_vp, _vp, C.POINTER(Beam), C.POINTER(Params), C.POINTER(GradOpts), _ip, C.c_int,
                             C.c_int, _dp, _dp, _dp, _dp, C.POINTER(Stats)]),
    "rte_fisher_inverse": (C.c_int, [C.c_int, _dp, _dp]),
    "rte_ref_jacobian": (C.c_int, [_vp, _vp, C.POINTER(Beam), C.POINTER(Params), _ip, C.c_int, C.c_int, _dp, _dp,
                                   C.POINTER(Stats)]),
    "rte_server_create": (C.c_int, [_vp, C.POINTER(ServerOpts), C.POINTER(_vp)]),
    "rte_server_destroy": (C.c_int, [_vp]),
    "rte_submit": (C.c_int, [_vp, C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam), C.POINTER(Params),
                             C.POINTER(C.c_longlong)]),
    "rte_poll": (C.c_int, [_vp, C.c_longlong, _dp, C.POINTER(Stats), _ip]),
    "rte_wait": (C.c_int, [_vp, C.c_longlong, C.c_double, _dp, C.POINTER(Stats), _ip]),
    "rte_cancel": (C.c_int, [_vp, C.c_longlong]),
    "rte_server_info": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "rte_propagate": (C.c_int, [_vp, C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_double,
                                C.c_int, C.c_int, _dp, _dp]),
    "rte_propagate_noalloc": (C.c_int, [_vp, C.POINTER(Grid), C.POINTER(Optics), C.c_int, C.c_double, C.c_int,
                                        C.c_int, _dp, _dp]),
    "rte_scatter": (C.c_int, [_vp, C.POINTER(Grid), C.POINTER(Optics), C.c_int, _dp]),
    "rte_source": (C.c_int, [_vp, C.POINTER(Grid), C.POINTER(Optics), C.POINTER(Beam), C.c_int,
                             _dp, _dp, _ip]),
    "rte_exitance": (C.c_int, [_vp, C.POINTER(Grid), C.c_int, C.c_int, C.c_double, _dp, _dp]),
    "rte_path_integrals": (C.c_int, [_vp, C.POINTER(Grid), C.POINTER(Optics), C.c_int, _ip, _ip, _dp]),
}

_LIB = None


def lib():
    """Loads librte.so; raises if it is missing (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librte.so not built: run `python -m doi_amd.build` (%s)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != OK:
        raise RteError(rc, lib().rte_last_error().decode(errors="replace"))
    return rc


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def fisher_inverse(chol):
    """F^-1 from the lower Cholesky factor rte_fisher returns (rte_fisher_inverse, FP64)."""
    chol = np.ascontiguousarray(chol, dtype=np.float64)
    n = chol.shape[0]
    out = np.zeros((n, n))
    _check(lib().rte_fisher_inverse(n, _p(chol), _p(out)))
    return out


def to_reference_layout(field, L):
    """Real-basis coefficient-major field [K, N] -> the reference's layout: complex
    coefficients of the complex harmonics (Y_l,-m = (-1)^m conj Y_lm), voxel-major
    [N, K] (GPU package/src/defines.h:59).  c_l0 = r_l0, c_lm = (r_lm - i r_l,-m)/sqrt2,
    c_l,-m = (-1)^m conj c_lm (m > 0); rte_neumann_ref does the same on the device."""
    K = (L + 1) ** 2
    r = np.asarray(field, dtype=np.float64).reshape(K, -1)
    c = np.zeros(r.shape, dtype=np.complex128)
    for l in range(L + 1):
        c[l * l + l] = r[l * l + l]
        for m in range(1, l + 1):
            cp = (r[l * l + l + m] - 1j * r[l * l + l - m]) / np.sqrt(2.0)
            c[l * l + l + m] = cp
            c[l * l + l - m] = (-1) ** m * np.conj(cp)
    return np.ascontiguousarray(c.T)


def from_reference_layout(cfield, L):
    """The inverse of to_reference_layout for the field of a real radiance:
    [N, K] complex -> [K, N] real (r_lm = sqrt2 Re c_lm, r_l,-m = -sqrt2 Im c_lm)."""
    K = (L + 1) ** 2
    c = np.asarray(cfield, dtype=np.complex128).reshape(-1, K).T
    r = np.zeros(c.shape)
    for l in range(L + 1):
        r[l * l + l] = c[l * l + l].real
        for m in range(1, l + 1):
            r[l * l + l + m] = np.sqrt(2.0) * c[l * l + l + m].real
            r[l * l + l - m] = -np.sqrt(2.0) * c[l * l + l + m].imag
    return r


def make_params(L, precision=PREC_MIXED, stop_rule=STOP_LAMBDA, tol=1e-6, max_terms=1000,
                prop_thresh=0.0, exit_quad=QUAD_REF, path_mode=PATH_CACHED):
    return Params(L, max_terms, stop_rule, tol, prop_thresh, precision, exit_quad, path_mode)


class Medium:
    """Optical maps of one grid, resident on the device."""

    def __init__(self, ctx, shape, bounds, mu_a, mu_s, g, n=1.0, c=2.9979e10):
        nx, ny, nz = shape
        self.ctx = ctx
        self.shape = (nx, ny, nz)
        self.N = nx * ny * nz
        self.grid = Grid(nx, ny, nz, *bounds)
        self.mu_a = _f64(np.broadcast_to(mu_a, (self.N,)) if np.ndim(mu_a) == 0 else np.ravel(mu_a))
        self.mu_s = _f64(np.broadcast_to(mu_s, (self.N,)) if np.ndim(mu_s) == 0 else np.ravel(mu_s))
        self.optics = Optics(_p(self.mu_a), _p(self.mu_s), g, n, c)
        self.h = _vp()
        _check(lib().rte_medium_create(ctx.h, C.byref(self.grid), C.byref(self.optics), C.byref(self.h)))
        ctx._media.add(self)

    def reset_cache(self):
        _check(lib().rte_medium_reset_cache(self.h))

    def close(self):
        if self.h:
            lib().rte_medium_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device; owns a HIP stream and cached per-shape tables."""

    def __init__(self, device=0):
        self.h = _vp()
        self._media = weakref.WeakSet()
        self._servers = weakref.WeakSet()
        _check(lib().rte_ctx_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            for s in list(self._servers):  # a server's worker uses the context
                s.close()
            for m in list(self._media):   # media die with their context
                m.close()
            lib().rte_ctx_destroy(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_group_size(self, max_rhs):
        _check(lib().rte_ctx_set_group_size(self.h, max_rhs))

    def profile(self, on=True):
        _check(lib().rte_ctx_profile(self.h, 1 if on else 0))

    def get_profile(self):
        p = Prof()
        _check(lib().rte_ctx_get_profile(self.h, C.byref(p)))
        out = {}
        for k, _ in Prof._fields_:
            v = getattr(p, k)
            out[k] = list(v) if k.startswith("group_") else v
        return out

    def medium(self, shape, bounds, mu_a, mu_s, g, n=1.0, c=2.9979e10):
        return Medium(self, shape, bounds, mu_a, mu_s, g, n, c)

    def solve_media(self, media, beams, params):
        nb = len(media)
        nx, ny, _ = media[0].shape
        arr = (_vp * nb)(*[m.h for m in media])
        barr = (Beam * nb)(*[Beam(*b) for b in beams])
        img = np.zeros(nb * nx * ny)
        st = (Stats * nb)()
        _check(lib().rte_solve_media(self.h, nb, arr, barr, C.byref(params), _p(img), st))
        return img.reshape(nb, ny, nx), [dict(terms=s.terms, lam=s.lambda_last, rel=s.rel_change_last,
                                              total=s.total_exitance, ballistic=s.ballistic_total,
                                              gpu_ms=s.gpu_ms, latency_ms=s.latency_ms) for s in st]

    def solve_images(self, media, beams, params):
        """(transmitted images, reflectance images, stats) of a batch (rte_solve_images)."""
        nb = len(media)
        nx, ny, _ = media[0].shape
        arr = (_vp * nb)(*[m.h for m in media])
        barr = (Beam * nb)(*[Beam(*b) for b in beams])
        img = np.zeros(nb * nx * ny)
        refl = np.zeros(nb * nx * ny)
        st = (Stats * nb)()
        _check(lib().rte_solve_images(self.h, nb, arr, barr, C.byref(params), _p(img), _p(refl), st))
        return img.reshape(nb, ny, nx), refl.reshape(nb, ny, nx), [_stats_dict(s) for s in st]

    def solve_field(self, medium, beam, params):
        """(final field [K, N] in the reference's units, transmitted image, reflectance, stats)."""
        nx, ny, _ = medium.shape
        K = (params.L + 1) ** 2
        field = np.zeros(K * medium.N)
        img = np.zeros(nx * ny)
        refl = np.zeros(nx * ny)
        st = Stats()
        _check(lib().rte_solve_field(self.h, medium.h, C.byref(Beam(*beam)), C.byref(params), _p(field), _p(img),
                                     _p(refl), C.byref(st)))
        return field.reshape(K, medium.N), img.reshape(ny, nx), refl.reshape(ny, nx), _stats_dict(st)

    def neumann(self, medium, src, params):
        """The reference's Neumann(src, out, 1) (rte_neumann): W_0 = src ([K, N], the
        reference's units), field = A sum_q W_q.  Returns (field [K, N], transmitted
        image, reflectance, stats), the images without a ballistic term."""
        nx, ny, _ = medium.shape
        K = (params.L + 1) ** 2
        src = _f64(np.reshape(src, (K * medium.N,)))
        field = np.zeros(K * medium.N)
        img = np.zeros(nx * ny)
        refl = np.zeros(nx * ny)
        st = Stats()
        _check(lib().rte_neumann(self.h, medium.h, C.byref(params), _p(src), _p(field), _p(img), _p(refl),
                                 C.byref(st)))
        return field.reshape(K, medium.N), img.reshape(ny, nx), refl.reshape(ny, nx), _stats_dict(st)

    def neumann_ref(self, medium, src, params, flag=1, out=None):
        """The reference's Neumann(src, out, flag) on its own layout (rte_neumann_ref):
        src and out complex [N, K] (voxel-major, complex SH basis, reference units).
        flag 1: out defaults to a copy of src, as the reference's callers preset it,
        and receives A sum_q (DA)^q src; flag 0: out <- A out.  Returns (out, stats)."""
        K = (params.L + 1) ** 2
        if flag == 1:
            src = np.ascontiguousarray(np.reshape(src, (medium.N, K)), dtype=np.complex128)
            out = src.copy() if out is None else np.ascontiguousarray(out, dtype=np.complex128).copy()
        else:
            out = np.ascontiguousarray(np.reshape(out, (medium.N, K)), dtype=np.complex128).copy()
        st = Stats()
        sp = src.ctypes.data_as(_dp) if flag == 1 else None
        _check(lib().rte_neumann_ref(self.h, medium.h, C.byref(params), sp, out.ctypes.data_as(_dp), flag,
                                     C.byref(st)))
        return out.reshape(medium.N, K), _stats_dict(st)

    def solve_fd(self, media, beams, omegas, params, field=False):
        """Frequency-domain solves (rte_solve_fd): complex transmitted and
        reflectance images [nb, ny, nx], stats, and with field=True the final
        complex field [K, N] of solve 0."""
        nb = len(media)
        nx, ny, _ = media[0].shape
        K = (params.L + 1) ** 2
        arr = (_vp * nb)(*[m.h for m in media])
        barr = (Beam * nb)(*[Beam(*b) for b in beams])
        om = _f64(np.broadcast_to(np.asarray(omegas, dtype=np.float64), (nb,)))
        img = np.zeros(nb * 2 * nx * ny)
        refl = np.zeros(nb * 2 * nx * ny)
        fld = np.zeros(2 * K * media[0].N) if field else None
        st = (Stats * nb)()
        _check(lib().rte_solve_fd(self.h, nb, arr, barr, _p(om), C.byref(params), _p(img), _p(refl),
                                  _p(fld) if field else None, st))
        img = img.reshape(nb, 2, ny, nx)
        refl = refl.reshape(nb, 2, ny, nx)
        out = (img[:, 0] + 1j * img[:, 1], refl[:, 0] + 1j * refl[:, 1], [_stats_dict(s) for s in st])
        if field:
            f = fld.reshape(2, K, media[0].N)
            out = out + (f[0] + 1j * f[1],)
        return out

    def solve(self, shape, bounds, mu_a, mu_s, g, beam, params, n=1.0, c=2.9979e10):
        m = self.medium(shape, bounds, mu_a, mu_s, g, n, c)
        try:
            img, st = self.solve_media([m], [beam], params)
        finally:
            m.close()
        return img[0], st[0]

    # ---- derivatives -----------------------------------------------------
    def exitance_vjp(self, medium, beam, params, cotangent, checkpoint_stride=1):
        nx, ny, _ = medium.shape
        cot = _f64(np.ravel(cotangent))
        img = np.zeros(nx * ny)
        ga = np.zeros(medium.N)
        gs = np.zeros(medium.N)
        st = Stats()
        _check(lib().rte_exitance_vjp(self.h, medium.h, C.byref(Beam(*beam)), C.byref(params),
                                      C.byref(GradOpts(checkpoint_stride, 0)), _p(cot), _p(img), _p(ga), _p(gs),
                                      C.byref(st)))
        return img.reshape(ny, nx), ga, gs, dict(terms=st.terms, gpu_ms=st.gpu_ms)

    def gradient(self, medium, beam, params, target, checkpoint_stride=1):
        nx, ny, _ = medium.shape
        tgt = _f64(np.ravel(target))
        img = np.zeros(nx * ny)
        ga = np.zeros(medium.N)
        gs = np.zeros(medium.N)
        loss = C.c_double(0.0)
        st = Stats()
        _check(lib().rte_gradient(self.h, medium.h, C.byref(Beam(*beam)), C.byref(params),
                                  C.byref(GradOpts(checkpoint_stride, 0)), _p(tgt), C.byref(loss), _p(img), _p(ga),
                                  _p(gs), C.byref(st)))
        return loss.value, img.reshape(ny, nx), ga, gs, dict(terms=st.terms, gpu_ms=st.gpu_ms)

    def fisher(self, medium, beam, params, partition, nparam, wrt=0, checkpoint_stride=1, jacobian=0):
        """jacobian 0: exact (adjoint per detector); 1: the reference's forward-mode derivative."""
        nx, ny, _ = medium.shape
        part = np.ascontiguousarray(np.ravel(partition), dtype=np.int32)
        F = np.zeros((nparam, nparam))
        Cf = np.zeros((nparam, nparam))
        J = np.zeros((nx * ny, nparam))
        img = np.zeros(nx * ny)
        st = Stats()
        _check(lib().rte_fisher(self.h, medium.h, C.byref(Beam(*beam)), C.byref(params),
                                C.byref(GradOpts(checkpoint_stride, jacobian)), part.ctypes.data_as(_ip), nparam,
                                wrt, _p(F), _p(Cf), _p(J), _p(img), C.byref(st)))
        return F, Cf, J, img.reshape(ny, nx), dict(terms=st.terms, gpu_ms=st.gpu_ms)

    def ref_jacobian(self, medium, beam, params, partition, nparam, wrt=0):
        """The reference's forward-mode derivative images, J[p, a] (rte_ref_jacobian)."""
        nx, ny, _ = medium.shape
        part = np.ascontiguousarray(np.ravel(partition), dtype=np.int32)
        J = np.zeros((nx * ny, nparam))
        img = np.zeros(nx * ny)
        st = Stats()
        _check(lib().rte_ref_jacobian(self.h, medium.h, C.byref(Beam(*beam)), C.byref(params),
                                      part.ctypes.data_as(_ip), nparam, wrt, _p(J), _p(img), C.byref(st)))
        return J, img.reshape(ny, nx), dict(terms=st.terms, gpu_ms=st.gpu_ms)

    # ---- operators -------------------------------------------------------
    def propagate(self, shape, bounds, mu_a, mu_s, g, W, L, prop_thresh=0.0, precision=PREC_FP64,
                  path_mode=PATH_CACHED, n=1.0, c=2.9979e10):
        m = _Maps(shape, bounds, mu_a, mu_s, g, n, c)
        W = _f64(W)
        out = np.zeros_like(W)
        _check(lib().rte_propagate(self.h, C.byref(m.grid), C.byref(m.optics), L, prop_thresh,
                                   precision, path_mode, _p(W), _p(out)))
        return out

    def propagate_noalloc(self, shape, bounds, mu_a, mu_s, g, W, L, prop_thresh=0.0, precision=PREC_FP64,
                          path_mode=PATH_MARCH, n=1.0, c=2.9979e10):
        """rte_propagate_noalloc: A W with every device buffer in the library's static
        scratch array (marched path integrals only)."""
        m = _Maps(shape, bounds, mu_a, mu_s, g, n, c)
        W = _f64(W)
        out = np.zeros_like(W)
        _check(lib().rte_propagate_noalloc(self.h, C.byref(m.grid), C.byref(m.optics), L, prop_thresh,
                                           precision, path_mode, _p(W), _p(out)))
        return out

    def scatter(self, shape, bounds, mu_a, mu_s, g, W, L, n=1.0, c=2.9979e10):
        m = _Maps(shape, bounds, mu_a, mu_s, g, n, c)
        W = _f64(W).copy()
        _check(lib().rte_scatter(self.h, C.byref(m.grid), C.byref(m.optics), L, _p(W)))
        return W

    def source(self, shape, bounds, mu_a, mu_s, g, beam, L, n=1.0, c=2.9979e10):
        m = _Maps(shape, bounds, mu_a, mu_s, g, n, c)
        K = (L + 1) ** 2
        s0 = np.zeros((K, m.N))
        ball = np.zeros(shape[0] * shape[1])
        ncols = C.c_int(0)
        _check(lib().rte_source(self.h, C.byref(m.grid), C.byref(m.optics), C.byref(Beam(*beam)), L,
                                _p(s0), _p(ball), C.byref(ncols)))
        return s0, ball, ncols.value

    def exitance(self, shape, bounds, w, L, quad=QUAD_REF, c_m=2.9979e10):
        grid = Grid(*shape, *bounds)
        img = np.zeros(shape[0] * shape[1])
        _check(lib().rte_exitance(self.h, C.byref(grid), L, quad, c_m, _p(_f64(w)), _p(img)))
        return img

    def path_integrals(self, shape, bounds, mu_a, mu_s, dest, src):
        m = _Maps(shape, bounds, mu_a, mu_s, 0.0, 1.0, 1.0)
        d = np.ascontiguousarray(dest, dtype=np.int32)
        s = np.ascontiguousarray(src, dtype=np.int32)
        tau = np.zeros(d.size)
        _check(lib().rte_path_integrals(self.h, C.byref(m.grid), C.byref(m.optics), d.size,
                                        d.ctypes.data_as(_ip), s.ctypes.data_as(_ip), _p(tau)))
        return tau


class _Maps:
    def __init__(self, shape, bounds, mu_a, mu_s, g, n, c):
        nx, ny, nz = shape
        self.N = nx * ny * nz
        self.grid = Grid(nx, ny, nz, *bounds)
        self.mu_a = _f64(np.broadcast_to(mu_a, (self.N,)) if np.ndim(mu_a) == 0 else np.ravel(mu_a))
        self.mu_s = _f64(np.broadcast_to(mu_s, (self.N,)) if np.ndim(mu_s) == 0 else np.ravel(mu_s))
        self.optics = Optics(_p(self.mu_a), _p(self.mu_s), g, n, c)


def footprint(shape, bounds, params, path_mode=PATH_CACHED):
    return lib().rte_solve_footprint(C.byref(Grid(*shape, *bounds)), C.byref(params), path_mode)


def _stats_dict(s):
    return dict(terms=s.terms, lam=s.lambda_last, rel=s.rel_change_last, total=s.total_exitance,
                ballistic=s.ballistic_total, gpu_ms=s.gpu_ms, latency_ms=s.latency_ms)


class Server:
    """Continuous-batching solve service on one context (rte_server_*).

    While it exists its worker thread owns the context: make no other calls on
    that context until close()."""

    def __init__(self, ctx, hbm_budget=0, max_slots_per_bucket=0, max_queue=0):
        self.ctx = ctx
        self.h = _vp()
        self._shapes = {}
        _check(lib().rte_server_create(ctx.h, C.byref(ServerOpts(hbm_budget, max_slots_per_bucket, max_queue)),
                                       C.byref(self.h)))
        ctx._servers.add(self)

    def submit(self, shape, bounds, mu_a, mu_s, g, beam, params, n=1.0, c=2.9979e10):
        nx, ny, nz = shape
        N = nx * ny * nz
        ma = _f64(np.broadcast_to(mu_a, (N,)) if np.ndim(mu_a) == 0 else np.ravel(mu_a))
        ms = _f64(np.broadcast_to(mu_s, (N,)) if np.ndim(mu_s) == 0 else np.ravel(mu_s))
        rid = C.c_longlong(0)
        _check(lib().rte_submit(self.h, C.byref(Grid(nx, ny, nz, *bounds)), C.byref(Optics(_p(ma), _p(ms), g, n, c)),
                                C.byref(Beam(*beam)), C.byref(params), C.byref(rid)))
        self._shapes[rid.value] = (ny, nx)
        return rid.value

    def _collect(self, rid, fn, *extra):
        ny, nx = self._shapes[rid]
        img = np.zeros(nx * ny)
        st = Stats()
        status = C.c_int(0)
        try:
            _check(fn(self.h, rid, *extra, _p(img), C.byref(st), C.byref(status)))
        except RteError as e:
            if e.code != E_BUSY:
                self._shapes.pop(rid, None)
            raise
        if status.value in (REQ_DONE, REQ_CANCELLED):
            del self._shapes[rid]
        if status.value == REQ_DONE:
            return status.value, img.reshape(ny, nx), _stats_dict(st)
        return status.value, None, None

    def poll(self, rid):
        """(status, image or None, stats or None); raises for a failed solve."""
        return self._collect(rid, lib().rte_poll)

    def wait(self, rid, timeout_ms=-1.0):
        """As poll, blocking until the solve leaves the service (or timeout_ms)."""
        return self._collect(rid, lib().rte_wait, timeout_ms)

    def cancel(self, rid):
        _check(lib().rte_cancel(self.h, rid))

    def info(self):
        s = ServerStats()
        _check(lib().rte_server_info(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().rte_server_destroy(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
