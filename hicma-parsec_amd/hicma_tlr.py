"""ctypes binding of libhicma_tlr.so, the MI355X TLR Cholesky library.

Mirrors the reference's C interface (hicma_parsec/hicma_parsec.h:38-211):
``STARSH_gen``, ``parsec_rank_gather``, ``parsec_band_size_auto_tuning``,
``parsec_band_regenerate``, ``HiCMA_dpotrf_L``, ``STARSH_check``.  The binding
only marshals arguments; every computation runs in the HIP library, and
loading fails loudly if the library is missing.
"""
import ctypes as C
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
    "HICMA_get_stat2", "parsec_band_size_auto_tuning", "hicma_band_autotune", "parsec_band_regenerate",
    "HiCMA_dpotrf_L", "hicma_op_counts", "hicma_op_stats", "hicma_rank_overflows", "STARSH_check",
    "hicma_get_tile", "hicma_set_tile", "hicma_get_rank_map", "hicma_logdet", "hicma_recompress_stats",
    "hicma_profile", "hicma_profile_read", "hicma_desc_copy",
]

_lib = None


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
            "STARSH_gen": (i, [vp, i, vp, vp, d, i, i, i, C.POINTER(i)]),
            "parsec_rank_gather": (i, [vp, vp, ip, i]),
            "parsec_rank_check": (i, [vp, vp, i]),
            "HICMA_get_stat2": (None, [ip, i, i, C.POINTER(i), C.POINTER(i), C.POINTER(d)]),
            "parsec_band_size_auto_tuning": (i, [vp, vp, vp, ip, i]),
            "hicma_band_autotune": (i, [ip, i, i]),
            "parsec_band_regenerate": (i, [vp, i, vp, i]),
            "HiCMA_dpotrf_L": (i, [vp, i, vp, vp, vp, d, i, i, C.POINTER(i), i, i, i, i, C.POINTER(i),
                                   vp, vp, vp, i]),
            "hicma_op_counts": (C.c_ulong, [C.c_char, C.c_ulong, C.c_ulong, C.c_ulong, C.c_ulong]),
            "hicma_op_stats": (i, [vp, vp]),
            "hicma_rank_overflows": (i, [vp]),
            "STARSH_check": (i, [vp, i, dp, vp, vp, i, i, C.POINTER(i)]),
            "hicma_get_tile": (i, [vp, i, i, dp, C.POINTER(i)]),
            "hicma_set_tile": (i, [vp, i, i, dp, i]),
            "hicma_get_rank_map": (i, [vp, ip]),
            "hicma_logdet": (i, [vp, C.POINTER(d)]),
            "hicma_recompress_stats": (i, [vp, vp]),
            "hicma_profile": (i, [vp, i]),
            "hicma_profile_read": (i, [vp, C.POINTER(d), C.POINTER(i)]),
            "hicma_desc_copy": (i, [vp, vp]),
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


def band_autotune(G, NT, NB):
    G = np.ascontiguousarray(G, dtype=np.int32)
    return lib().hicma_band_autotune(G.ctypes.data, NT, NB)


def rank_stat(G, NT, band):
    G = np.ascontiguousarray(G, dtype=np.int32)
    mn, mx, av = C.c_int(), C.c_int(), C.c_double()
    lib().HICMA_get_stat2(G.ctypes.data, NT, band, C.byref(mn), C.byref(mx), C.byref(av))
    return mn.value, mx.value, av.value


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().hicma_init(C.byref(self.h), device), "hicma_init")

    def close(self):
        if self.h:
            lib().hicma_fini(self.h)
            self.h = C.c_void_p()

    def profile(self, enable=True):
        _check(lib().hicma_profile(self.h, int(enable)), "hicma_profile")

    def profile_read(self):
        ms, n = C.c_double(), C.c_int()
        _check(lib().hicma_profile_read(self.h, C.byref(ms), C.byref(n)), "hicma_profile_read")
        return ms.value, n.value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class TLRMatrix:
    """A TLR descriptor: dense band tiles, [U|V] off-band tiles, rank map (HBM)."""

    def __init__(self, ctx, N, NB, maxrank, kind=SQEXP_2D, sigma2=1.0, beta=0.1, nugget=1e-2, seed=0):
        self.ctx, self.N, self.NB, self.maxrank, self.kind = ctx, N, NB, maxrank, kind
        self.NT = (N + NB - 1) // NB
        self.h = C.c_void_p()
        _check(lib().hicma_desc_create(ctx.h, C.byref(self.h), N, NB, maxrank, kind, sigma2, beta, nugget, seed),
               "hicma_desc_create")
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

    def generate(self, tol=1e-8, maxrank=None, band_size=1):
        info = C.c_int(0)
        _check(lib().STARSH_gen(self.ctx.h, LOWER, self.h, self.h, tol, maxrank or self.maxrank, band_size, 0,
                                C.byref(info)), "STARSH_gen")
        self.band = band_size
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

    def regenerate(self, band_size):
        _check(lib().parsec_band_regenerate(self.ctx.h, LOWER, self.h, band_size), "parsec_band_regenerate")
        self.band = band_size

    def potrf(self, acc=1e-8, fixed_rank=0, maxrank=None, lookahead=-1, two_flow=0, want_ops=False):
        la, tf = C.c_int(lookahead), C.c_int(two_flow)
        tile = np.zeros(self.NT * self.NT, dtype=np.uint64) if want_ops else None
        opc = np.zeros(4, dtype=np.uint64)
        cpt = np.zeros(6)
        mr = maxrank or self.maxrank
        info = lib().HiCMA_dpotrf_L(self.ctx.h, LOWER, self.h, self.h, self.h, acc, fixed_rank, mr, C.byref(la),
                                    self.band, 300, mr, 0, C.byref(tf),
                                    tile.ctypes.data if want_ops else None, opc.ctypes.data, cpt.ctypes.data, 0)
        _check(info, "HiCMA_dpotrf_L")
        st = np.zeros(4, dtype=np.uint64)
        lib().hicma_op_stats(self.h, st.ctypes.data)
        self.stats = st
        self.two_flow = tf.value
        self.critical_path_time = cpt
        self.tileops = tile.reshape(self.NT, self.NT) if want_ops else None
        return info

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

    def uncompress(self):
        D = np.zeros((self.N, self.N), order="F")
        info = C.c_int(0)
        _check(lib().STARSH_check(self.ctx.h, LOWER, D.ctypes.data, self.h, self.h, self.band, self.maxrank,
                                  C.byref(info)), "STARSH_check")
        return D

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
