"""ctypes binding of libhicma_tlr.so, the MI355X TLR Cholesky library.

Mirrors the reference's C interface (hicma_parsec/hicma_parsec.h:38-211):
``STARSH_gen``, ``parsec_rank_gather``, ``parsec_band_size_auto_tuning``,
``parsec_band_regenerate``, ``HiCMA_dpotrf_L``, ``STARSH_check``, and the
harness's checks (``check_dpotrf2``, ``check_diff``, hicma_parsec/testing_dpotrf.c:111-230).  The binding
only marshals arguments; every computation runs in the HIP library, and
loading fails loudly if the library is missing.
"""
import ctypes as C
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libhicma_tlr.so")

LOWER, UPPER = 122, 121
SQEXP_2D, SQEXP_3D, EXP_3D = 2, 3, 4
ERR = {-1: "uplo must be Lower", -2: "maxrank > NB/2", -3: "maxrank > compmaxrank",
       -4: "out of HBM", -5: "HIP runtime error", -6: "invalid argument"}

# every entry point declared in include/hicma_tlr.h
EXPORTS = [
    "hicma_init", "hicma_fini", "hicma_desc_create", "hicma_desc_destroy", "hicma_desc_info",
    "hicma_get_points", "hicma_points", "STARSH_gen", "parsec_rank_gather", "parsec_rank_check",
    "HICMA_get_stat2", "parsec_band_size_auto_tuning", "hicma_band_autotune", "hicma_band_autotune_weighted",
    "hicma_band_autotune_model",
    "parsec_band_regenerate",
    "HiCMA_dpotrf_L", "hicma_op_counts", "hicma_op_stats", "hicma_rank_overflows", "STARSH_check",
    "hicma_get_tile", "hicma_set_tile", "hicma_get_rank_map", "hicma_logdet", "hicma_recompress_stats",
    "hicma_profile", "hicma_profile_read", "hicma_desc_copy", "hicma_jacobi_unconverged", "hicma_exec_op_stats",
    "hicma_rank_history", "hicma_get_rank_history", "parsec_rank_print", "hicma_dense_create", "hicma_dense_destroy", "hicma_dense_get",
    "hicma_dense_set", "hicma_dense_generate", "hicma_dense_potrf", "hicma_dense_lansy", "hicma_dense_diff_lansy",
    "hicma_dense_logdet", "hicma_check_dpotrf2", "hicma_check_diff", "hicma_tlr_trsm", "hicma_tlr_potrs",
    "hicma_tlr_loglik", "hicma_dev_malloc", "hicma_dev_free", "hicma_memcpy", "hicma_tlr_loglik_noalloc",
    "hicma_init_ex", "hicma_server_create", "hicma_server_submit", "hicma_server_wait", "hicma_server_destroy",
    "hicma_normal", "hicma_factor_graph", "hicma_set_wave_min_tasks", "hicma_autotune_schedule",
    "hicma_tlr_normf", "hicma_tlr_demote", "hicma_gen_first", "hicma_gen_finish",
]

_lib = None


class LikRequest(C.Structure):
    """hicma_lik_request"""
    _fields_ = [("N", C.c_int), ("NB", C.c_int), ("maxrank", C.c_int), ("eps", C.c_double),
                ("sigma2", C.c_double), ("beta", C.c_double), ("nugget", C.c_double), ("seed", C.c_uint64),
                ("z", C.c_void_p), ("zseed", C.c_uint64)]


class LikResult(C.Structure):
    """hicma_lik_result"""
    _fields_ = [("status", C.c_int), ("band", C.c_int), ("logdet", C.c_double), ("quad", C.c_double),
                ("loglik", C.c_double), ("t_queue", C.c_double), ("t_run", C.c_double), ("t_total", C.c_double),
                ("bytes", C.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        vp, i, d, u64, ip, dp = C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_void_p, C.c_void_p
        sig = {
            "hicma_init": (i, [C.POINTER(vp), i]),
            "hicma_fini": (i, [vp]),
            "hicma_desc_create": (i, [vp, C.POINTER(vp), i, i, i, i, d, d, d, u64]),
            "hicma_desc_destroy": (i, [vp]),
            "hicma_desc_info": (i, [vp, ip, ip, ip, ip, ip]),
            "hicma_get_points": (i, [vp, dp]),
            "hicma_points": (i, [i, i, u64, dp]),
            "STARSH_gen": (i, [vp, i, vp, vp, vp, vp, vp, d, i, i, i, C.POINTER(i)]),
            "parsec_rank_gather": (i, [vp, vp, ip, i]),
            "parsec_rank_check": (i, [vp, vp, i]),
            "HICMA_get_stat2": (None, [ip, i, i, C.POINTER(i), C.POINTER(i), C.POINTER(d)]),
            "parsec_band_size_auto_tuning": (i, [vp, vp, vp, ip, i]),
            "hicma_band_autotune": (i, [ip, i, i]),
            "hicma_band_autotune_weighted": (i, [ip, i, i, d]),
            "hicma_band_autotune_model": (i, [ip, i, i, d, d]),
            "parsec_band_regenerate": (i, [vp, i, vp, vp, vp, vp, i]),
            "HiCMA_dpotrf_L": (i, [vp, i, vp, vp, vp, d, i, i, C.POINTER(i), i, i, i, i, C.POINTER(i),
                                   vp, vp, vp, i]),
            "hicma_op_counts": (C.c_ulong, [C.c_char, C.c_ulong, C.c_ulong, C.c_ulong, C.c_ulong]),
            "hicma_op_stats": (i, [vp, vp]),
            "hicma_rank_overflows": (i, [vp]),
            "STARSH_check": (i, [vp, i, vp, vp, vp, i, i, C.POINTER(i)]),
            "hicma_get_tile": (i, [vp, i, i, dp, C.POINTER(i)]),
            "hicma_set_tile": (i, [vp, i, i, dp, i]),
            "hicma_get_rank_map": (i, [vp, ip]),
            "hicma_logdet": (i, [vp, C.POINTER(d)]),
            "hicma_recompress_stats": (i, [vp, vp]),
            "hicma_profile": (i, [vp, i]),
            "hicma_profile_read": (i, [vp, C.POINTER(d), C.POINTER(i)]),
            "hicma_desc_copy": (i, [vp, vp]),
            "hicma_jacobi_unconverged": (i, [vp]),
            "hicma_exec_op_stats": (i, [vp, vp, i]),
            "hicma_rank_history": (i, [vp, i]),
            "hicma_get_rank_history": (i, [vp, vp]),
            "parsec_rank_print": (i, [vp, vp, i]),
            "hicma_dense_create": (i, [vp, C.POINTER(vp), i]),
            "hicma_dense_destroy": (i, [vp]),
            "hicma_dense_get": (i, [vp, dp]),
            "hicma_dense_set": (i, [vp, dp]),
            "hicma_dense_generate": (i, [vp, i, vp, vp]),
            "hicma_dense_potrf": (i, [vp, i, vp, C.POINTER(i)]),
            "hicma_dense_lansy": (i, [vp, i, vp, C.POINTER(d)]),
            "hicma_dense_diff_lansy": (i, [vp, i, vp, vp, C.POINTER(d)]),
            "hicma_dense_logdet": (i, [vp, vp, C.POINTER(d)]),
            "hicma_check_dpotrf2": (i, [vp, i, i, vp, vp, d, C.POINTER(d)]),
            "hicma_check_diff": (i, [vp, i, i, vp, vp, d, C.POINTER(d)]),
            "hicma_tlr_trsm": (i, [vp, vp, i, i, vp, i]),
            "hicma_tlr_potrs": (i, [vp, vp, i, vp, i]),
            "hicma_tlr_loglik": (i, [vp, vp, vp, C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
            "hicma_dev_malloc": (i, [vp, C.c_size_t, C.POINTER(vp)]),
            "hicma_dev_free": (i, [vp, vp]),
            "hicma_memcpy": (i, [vp, vp, vp, C.c_size_t, i]),
            "hicma_tlr_loglik_noalloc": (i, [vp, vp, vp, C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
            "hicma_init_ex": (i, [C.POINTER(vp), i, i, i]),
            "hicma_server_create": (i, [C.POINTER(vp), i, i, d]),
            "hicma_server_submit": (i, [vp, C.POINTER(LikRequest), C.POINTER(C.c_uint64)]),
            "hicma_server_wait": (i, [vp, C.c_uint64, C.POINTER(LikResult)]),
            "hicma_server_destroy": (i, [vp]),
            "hicma_normal": (i, [u64, i, vp]),
            "hicma_factor_graph": (i, [vp, i]),
            "hicma_set_wave_min_tasks": (i, [vp, i]),
            "hicma_autotune_schedule": (i, [vp, vp, d, i, i, C.POINTER(i), C.POINTER(d)]),
            "hicma_tlr_normf": (i, [vp, C.POINTER(d), dp]),
            "hicma_tlr_demote": (i, [vp, d, d, ip]),
            "hicma_gen_first": (i, [vp, vp, d, i, C.POINTER(i)]),
            "hicma_gen_finish": (i, [vp, vp, d, i, C.POINTER(i)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc < 0:
        raise RuntimeError(f"{what} failed: {rc} ({ERR.get(rc, 'error')})")
    return rc


def points(N, ndim, seed=0):
    xyz = np.zeros(ndim * N)
    _check(lib().hicma_points(N, ndim, seed, xyz.ctypes.data), "hicma_points")
    return xyz.reshape(ndim, N)


def op_counts(op, a, b=0, c=0, d=0):
    return lib().hicma_op_counts(op.encode() if isinstance(op, str) else op, a, b, c, d)


def band_autotune(G, NT, NB, lr_weight=None):
    """hicma_band_autotune (the reference's rule), or with the low-rank op
    counts weighted (hicma_band_autotune_weighted)."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    if lr_weight is None:
        return lib().hicma_band_autotune(G.ctypes.data, NT, NB)
    return lib().hicma_band_autotune_weighted(G.ctypes.data, NT, NB, float(lr_weight))


def band_autotune_model(G, NT, NB, dense_tflops=10.0, lr_tflops=3.5):
    """hicma_band_autotune_model: the band with the least estimated time."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    return lib().hicma_band_autotune_model(G.ctypes.data, NT, NB, float(dense_tflops), float(lr_tflops))


# device rates of the serving loop's band model (TFLOP/s)
SERVE_DENSE_TFLOPS = 60.0
SERVE_PENDING_RANK = 160
SERVE_BAND_BYTES = 4 << 30
RANK_PENDING = -2
SERVE_LR_TFLOPS = 3.5


def rank_stat(G, NT, band):
    G = np.ascontiguousarray(G, dtype=np.int32)
    mn, mx, av = C.c_int(), C.c_int(), C.c_double()
    lib().HICMA_get_stat2(G.ctypes.data, NT, band, C.byref(mn), C.byref(mx), C.byref(av))
    return mn.value, mx.value, av.value


def normal(seed, n):
    z = np.zeros(n)
    _check(lib().hicma_normal(seed, n, z.ctypes.data), "hicma_normal")
    return z


class Server:
    """Continuous batching of Gaussian log-likelihood requests (hicma_server_*)."""

    def __init__(self, device=0, slots=4, hbm_budget_gb=200.0):
        self.h = C.c_void_p()
        _check(lib().hicma_server_create(C.byref(self.h), device, slots, hbm_budget_gb), "hicma_server_create")

    def submit(self, N, sigma2, beta, nugget=1e-2, seed=0, z=None, zseed=0, NB=512, maxrank=0, eps=1e-8):
        zz = None if z is None else np.ascontiguousarray(z, dtype=np.float64)
        r = LikRequest(N, NB, maxrank, eps, sigma2, beta, nugget, seed, None if zz is None else zz.ctypes.data, zseed)
        t = C.c_uint64()
        _check(lib().hicma_server_submit(self.h, C.byref(r), C.byref(t)), "hicma_server_submit")
        return t.value

    def wait(self, ticket):
        res = LikResult()
        _check(lib().hicma_server_wait(self.h, ticket, C.byref(res)), "hicma_server_wait")
        return res.as_dict()

    def close(self):
        if self.h:
            lib().hicma_server_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        self._owned = weakref.WeakSet()      # live objects holding device memory of this context
        _check(lib().hicma_init(C.byref(self.h), device), "hicma_init")

    def _own(self, obj):
        self._owned.add(obj)

    def close(self):
        """Frees every object still open on this context, then the context
        (an object kept alive past its context, e.g. by a traceback, must not
        free memory through a context that no longer exists)."""
        if self.h:
            for obj in list(self._owned):
                obj.close()
            lib().hicma_fini(self.h)
            self.h = C.c_void_p()

    def profile(self, enable=True):
        _check(lib().hicma_profile(self.h, int(enable)), "hicma_profile")

    def set_wave_min_tasks(self, ntask):
        _check(lib().hicma_set_wave_min_tasks(self.h, int(ntask)), "hicma_set_wave_min_tasks")

    def profile_read(self):
        ms, n = C.c_double(), C.c_int()
        _check(lib().hicma_profile_read(self.h, C.byref(ms), C.byref(n)), "hicma_profile_read")
        return ms.value, n.value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceArray:
    """A float64 array in HBM (N x k column major), for the solves."""

    def __init__(self, ctx, host):
        self.ctx = ctx
        self.host = np.array(host, dtype=np.float64, order="F", copy=True)
        self.p = C.c_void_p()
        _check(lib().hicma_dev_malloc(ctx.h, self.host.nbytes, C.byref(self.p)), "hicma_dev_malloc")
        ctx._own(self)
        _check(lib().hicma_memcpy(ctx.h, self.p, self.host.ctypes.data, self.host.nbytes, 1), "hicma_memcpy")

    def get(self):
        out = np.empty_like(self.host, order="F")
        _check(lib().hicma_memcpy(self.ctx.h, out.ctypes.data, self.p, out.nbytes, 0), "hicma_memcpy")
        return out

    def close(self):
        if self.p:
            lib().hicma_dev_free(self.ctx.h, self.p)
            self.p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KernelParams(C.Structure):
    """hicma_kernel_params: the `data` argument of STARSH_gen."""
    _fields_ = [("sigma2", C.c_double), ("beta", C.c_double), ("nugget", C.c_double)]


def _problem_args(params, kernel, index, N):
    data = C.byref(KernelParams(*params)) if params is not None else None
    kern = C.byref(C.c_int(kernel)) if kernel is not None else None
    idx = None
    if index is not None:
        idx = np.ascontiguousarray(index, dtype=np.int64)
        assert idx.shape == (N,)
    return data, kern, idx


class Dense:
    """An N x N column-major matrix in HBM (the --check's dense descriptors)."""

    def __init__(self, ctx, N):
        self.ctx, self.N = ctx, N
        self.h = C.c_void_p()
        _check(lib().hicma_dense_create(ctx.h, C.byref(self.h), N), "hicma_dense_create")
        ctx._own(self)

    def close(self):
        if self.h:
            lib().hicma_dense_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self):
        D = np.zeros((self.N, self.N), order="F")
        _check(lib().hicma_dense_get(self.h, D.ctypes.data), "hicma_dense_get")
        return D

    def set(self, D):
        D = np.asfortranarray(D, dtype=np.float64)
        _check(lib().hicma_dense_set(self.h, D.ctypes.data), "hicma_dense_set")

    def generate(self, problem):
        _check(lib().hicma_dense_generate(self.ctx.h, LOWER, self.h, problem.h), "hicma_dense_generate")

    def uncompress(self, A):
        info = C.c_int(0)
        _check(lib().STARSH_check(self.ctx.h, LOWER, self.h, A.h, A.h, A.band, A.maxrank, C.byref(info)),
               "STARSH_check")

    def potrf(self):
        info = C.c_int(0)
        _check(lib().hicma_dense_potrf(self.ctx.h, LOWER, self.h, C.byref(info)), "hicma_dense_potrf")
        return info.value

    def lansy(self, other=None):
        v = C.c_double()
        if other is None:
            _check(lib().hicma_dense_lansy(self.ctx.h, LOWER, self.h, C.byref(v)), "hicma_dense_lansy")
        else:
            _check(lib().hicma_dense_diff_lansy(self.ctx.h, LOWER, self.h, other.h, C.byref(v)),
                   "hicma_dense_diff_lansy")
        return v.value

    def logdet(self):
        v = C.c_double()
        _check(lib().hicma_dense_logdet(self.ctx.h, self.h, C.byref(v)), "hicma_dense_logdet")
        return v.value


def check_dpotrf2(ctx, L, A0, threshold, verbose=0):
    """(info, ||A0 - L L^T|| / ||A0||) on the device."""
    r = C.c_double()
    bad = _check(lib().hicma_check_dpotrf2(ctx.h, verbose, LOWER, L.h, A0.h, threshold, C.byref(r)),
                 "hicma_check_dpotrf2")
    return bad, r.value


def check_diff(ctx, L, L0, threshold, verbose=0):
    """(info, ||L - L0|| / ||L||) on the device."""
    r = C.c_double()
    bad = _check(lib().hicma_check_diff(ctx.h, verbose, LOWER, L.h, L0.h, threshold, C.byref(r)), "hicma_check_diff")
    return bad, r.value


class TLRMatrix:
    """A TLR descriptor: dense band tiles, [U|V] off-band tiles, rank map (HBM)."""

    def __init__(self, ctx, N, NB, maxrank, kind=SQEXP_2D, sigma2=1.0, beta=0.1, nugget=1e-2, seed=0):
        self.ctx, self.N, self.NB, self.maxrank, self.kind = ctx, N, NB, maxrank, kind
        self.NT = (N + NB - 1) // NB
        self.h = C.c_void_p()
        _check(lib().hicma_desc_create(ctx.h, C.byref(self.h), N, NB, maxrank, kind, sigma2, beta, nugget, seed),
               "hicma_desc_create")
        ctx._own(self)
        self.band = 1
        self.stats = None

    def close(self):
        if self.h:
            lib().hicma_desc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ndim(self):
        return 2 if self.kind == SQEXP_2D else 3

    def points(self):
        xyz = np.zeros(self.ndim * self.N)
        _check(lib().hicma_get_points(self.h, xyz.ctypes.data), "hicma_get_points")
        return xyz.reshape(self.ndim, self.N)

    def generate(self, tol=1e-8, maxrank=None, band_size=1, params=None, kernel=None, index=None):
        """STARSH_gen; params = (sigma2, beta, nugget), kernel = a kind,
        index = point index of each matrix row (None: keep the descriptor's)."""
        info = C.c_int(0)
        data, kern, idx = _problem_args(params, kernel, index, self.N)
        _check(lib().STARSH_gen(self.ctx.h, LOWER, self.h, self.h, data, kern,
                                idx.ctypes.data if idx is not None else None, tol, maxrank or self.maxrank,
                                band_size, 0, C.byref(info)), "STARSH_gen")
        self.band = band_size
        return info.value

    def generate_first(self, tol=1e-8, maxrank=None):
        """hicma_gen_first: band 1, narrow sketch only (pending tiles at RANK_PENDING)."""
        info = C.c_int(0)
        _check(lib().hicma_gen_first(self.ctx.h, self.h, tol, maxrank or self.maxrank, C.byref(info)),
               "hicma_gen_first")
        self.band = 1
        return info.value

    def generate_finish(self, tol=1e-8, maxrank=None):
        """hicma_gen_finish: the wider sketches for the pending tiles outside the band."""
        info = C.c_int(0)
        _check(lib().hicma_gen_finish(self.ctx.h, self.h, tol, maxrank or self.maxrank, C.byref(info)),
               "hicma_gen_finish")
        return info.value

    def rank_gather(self, band=None):
        G = np.zeros(self.NT * self.NT, dtype=np.int32)
        _check(lib().parsec_rank_gather(self.ctx.h, self.h, G.ctypes.data, self.band if band is None else band),
               "parsec_rank_gather")
        return G

    def rank_map(self):
        G = np.zeros(self.NT * self.NT, dtype=np.int32)
        _check(lib().hicma_get_rank_map(self.h, G.ctypes.data), "hicma_get_rank_map")
        return G.reshape(self.NT, self.NT).T  # [m, n]

    def auto_tune(self):
        G = self.rank_gather()
        return _check(lib().parsec_band_size_auto_tuning(self.ctx.h, self.h, None, G.ctypes.data, self.NB),
                      "parsec_band_size_auto_tuning")

    def regenerate(self, band_size, params=None, kernel=None, index=None):
        data, kern, idx = _problem_args(params, kernel, index, self.N)
        _check(lib().parsec_band_regenerate(self.ctx.h, LOWER, self.h, data, kern,
                                            idx.ctypes.data if idx is not None else None, band_size),
               "parsec_band_regenerate")
        self.band = band_size

    def potrf(self, acc=1e-8, fixed_rank=0, maxrank=None, lookahead=-1, two_flow=0, want_ops=False,
              critical_path=False):
        la, tf = C.c_int(lookahead), C.c_int(two_flow)
        tile = np.zeros(self.NT * self.NT, dtype=np.uint64) if want_ops else None
        opc = np.zeros(4, dtype=np.uint64)
        cpt = np.zeros(6) if critical_path else None
        mr = maxrank or self.maxrank
        info = lib().HiCMA_dpotrf_L(self.ctx.h, LOWER, self.h, self.h, self.h, acc, fixed_rank, mr, C.byref(la),
                                    self.band, 300, mr, 0, C.byref(tf),
                                    tile.ctypes.data if want_ops else None, opc.ctypes.data,
                                    cpt.ctypes.data if critical_path else None, 0)
        _check(info, "HiCMA_dpotrf_L")
        st = np.zeros(4, dtype=np.uint64)
        lib().hicma_op_stats(self.h, st.ctypes.data)
        self.stats = st
        ex = np.zeros(4, dtype=np.uint64)
        lib().hicma_exec_op_stats(self.h, ex.ctypes.data, 4)
        self.exec_stats = ex
        self.two_flow = tf.value
        self.critical_path_time = cpt
        self.tileops = tile.reshape(self.NT, self.NT) if want_ops else None
        return info

    def trsm(self, B, trans=0):
        """L^-1 B (trans 0) or L^-T B (trans 1) on the device; B host array N x k or N."""
        X = DeviceArray(self.ctx, np.reshape(B, (self.N, -1), order="F"))
        _check(lib().hicma_tlr_trsm(self.ctx.h, self.h, trans, X.host.shape[1], X.p, self.N), "hicma_tlr_trsm")
        out = X.get()
        X.close()
        return out.reshape(np.shape(B), order="F")

    def potrs(self, B):
        X = DeviceArray(self.ctx, np.reshape(B, (self.N, -1), order="F"))
        _check(lib().hicma_tlr_potrs(self.ctx.h, self.h, X.host.shape[1], X.p, self.N), "hicma_tlr_potrs")
        out = X.get()
        X.close()
        return out.reshape(np.shape(B), order="F")

    def loglik(self, z):
        """(logdet, z^T Sigma^-1 z, log-likelihood) of the host vector z."""
        X = DeviceArray(self.ctx, np.asarray(z, dtype=np.float64).reshape(self.N, 1))
        ld, qd, ll = C.c_double(), C.c_double(), C.c_double()
        _check(lib().hicma_tlr_loglik(self.ctx.h, self.h, X.p, C.byref(ld), C.byref(qd), C.byref(ll)),
               "hicma_tlr_loglik")
        X.close()
        return ld.value, qd.value, ll.value

    def autotune_schedule(self, acc=1e-8, maxrank=None, reps=2):
        """(threshold, seconds) chosen by hicma_autotune_schedule; kept in the context."""
        b, sec = C.c_int(), C.c_double()
        _check(lib().hicma_autotune_schedule(self.ctx.h, self.h, acc, maxrank or self.maxrank, reps, C.byref(b),
                                             C.byref(sec)), "hicma_autotune_schedule")
        return b.value, sec.value

    def use_graph(self, enable=True):
        """Capture the factorization in a HIP graph once, replay it after."""
        _check(lib().hicma_factor_graph(self.h, int(enable)), "hicma_factor_graph")

    def jacobi_unconverged(self):
        return lib().hicma_jacobi_unconverged(self.h)

    def rank_history(self, enable=True):
        _check(lib().hicma_rank_history(self.h, int(enable)), "hicma_rank_history")

    def get_rank_history(self):
        """[m, n, q] with q = initial, min, max, final rank of the last factorization."""
        h = np.zeros(4 * self.NT * self.NT, dtype=np.int32)
        _check(lib().hicma_get_rank_history(self.h, h.ctypes.data), "hicma_get_rank_history")
        return h.reshape(self.NT, self.NT, 4).transpose(1, 0, 2)

    def rank_print(self):
        _check(lib().parsec_rank_print(self.ctx.h, self.h, self.band), "parsec_rank_print")

    def rank_check(self):
        """parsec_rank_check: number of negative off-band ranks; band entries set to -1."""
        return _check(lib().parsec_rank_check(self.ctx.h, self.h, self.band), "parsec_rank_check")

    def recompress_stats(self):
        st = np.zeros(3, dtype=np.uint64)
        _check(lib().hicma_recompress_stats(self.h, st.ctypes.data), "hicma_recompress_stats")
        return int(st[0]), int(st[1]), int(st[2])

    def copy_from(self, src):
        _check(lib().hicma_desc_copy(self.h, src.h), "hicma_desc_copy")
        self.band = src.band

    def rank_overflows(self):
        return lib().hicma_rank_overflows(self.h)

    def logdet(self):
        v = C.c_double()
        _check(lib().hicma_logdet(self.h, C.byref(v)), "hicma_logdet")
        return v.value

    def normf(self, tiles=False):
        """||A||_F of the symmetric matrix held (and the lower tiles' norms, NT x NT)."""
        v = C.c_double()
        T = np.zeros(self.NT * self.NT) if tiles else None
        _check(lib().hicma_tlr_normf(self.h, C.byref(v), T.ctypes.data if tiles else None), "hicma_tlr_normf")
        return (v.value, T.reshape(self.NT, self.NT, order="F")) if tiles else v.value

    def demote(self, thr32, thr16=0.0):
        """Hold off-band tiles with norm < thr32 in FP32, < thr16 in BF16;
        returns the tile counts (FP64, FP32, BF16)."""
        c = np.zeros(3, dtype=np.int32)
        _check(lib().hicma_tlr_demote(self.h, thr32, thr16, c.ctypes.data), "hicma_tlr_demote")
        return tuple(int(x) for x in c)

    def uncompress(self):
        """Host copy of the uncompressed lower triangle (STARSH_check into a
        device matrix, then copied back)."""
        D = Dense(self.ctx, self.N)
        try:
            D.uncompress(self)
            return D.get()
        finally:
            D.close()

    def get_tile(self, m, n):
        buf = np.zeros(self.NB * max(self.NB, 2 * self.maxrank))
        r = C.c_int()
        _check(lib().hicma_get_tile(self.h, m, n, buf.ctypes.data, C.byref(r)), "hicma_get_tile")
        if m - n < self.band:
            return buf[: self.NB * self.NB].reshape(self.NB, self.NB, order="F"), self.NB
        rk = r.value
        U = buf[: self.NB * rk].reshape(self.NB, rk, order="F")
        V = buf[self.NB * rk: 2 * self.NB * rk].reshape(self.NB, rk, order="F")
        return (U, V), rk

    def set_tile(self, m, n, data, rank):
        data = np.ascontiguousarray(data, dtype=np.float64)
        _check(lib().hicma_set_tile(self.h, m, n, data.ctypes.data, rank), "hicma_set_tile")
