"""ctypes binding of the CPU oracle (libtlr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / CPU reference, never as
the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libtlr_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} not built (make -C oracle)")
        L = C.CDLL(LIBPATH)
        vp, i, d, u64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64
        sig = {
            "orc_points": (i, [i, i, u64, vp]),
            "orc_sketch": (None, [i, i, u64, vp]),
            "orc_kernel_block": (None, [i, i, i, vp, i, i, i, d, d, d, vp, i]),
            "orc_create": (vp, [i, i, i, d, d, d, u64]),
            "orc_free": (None, [vp]),
            "orc_NT": (i, [vp]),
            "orc_xyz": (vp, [vp]),
            "orc_generate": (i, [vp, d, i, i]),
            "orc_band_regenerate": (i, [vp, i]),
            "orc_band_size": (i, [vp]),
            "orc_rank_gather": (None, [vp, vp]),
            "orc_rank_stat": (None, [vp, i, i, C.POINTER(i), C.POINTER(i), C.POINTER(d)]),
            "orc_rank_check": (i, [vp, i, i]),
            "orc_band_autotune": (i, [vp, i, i]),
            "orc_op_counts": (C.c_ulong, [C.c_char, C.c_ulong, C.c_ulong, C.c_ulong, C.c_ulong]),
            "orc_dpotrf": (i, [vp, d, i, i, i, vp, vp, C.POINTER(d)]),
            "orc_uncompress": (None, [vp, vp]),
            "orc_dense_exact": (None, [vp, vp]),
            "orc_logdet": (d, [vp]),
            "orc_check_dpotrf2": (d, [i, vp, vp]),
            "orc_check_diff": (d, [i, vp, vp]),
            "orc_dense_potrf": (i, [i, vp]),
            "orc_dense_logdet": (d, [i, vp]),
            "orc_get_rank": (i, [vp, i, i]),
            "orc_get_tile": (i, [vp, i, i, vp]),
            "orc_set_band": (i, [vp, i]),
            "orc_set_tile": (i, [vp, i, i, vp, i]),
            "orc_set_threads": (i, [i]),
            "orc_tlr_trsm": (i, [vp, i, i, vp, i]),
            "orc_blas_version": (None, [C.c_char_p, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def set_threads(n):
    return lib().orc_set_threads(n)


def blas_version():
    buf = C.create_string_buffer(256)
    lib().orc_blas_version(buf, 256)
    return buf.value.decode(errors="replace").strip()


def points(N, ndim, seed=0):
    xyz = np.zeros(ndim * N)
    lib().orc_points(N, ndim, seed, xyz.ctypes.data)
    return xyz.reshape(ndim, N)


def op_counts(op, a, b=0, c=0, d=0):
    return lib().orc_op_counts(op.encode(), a, b, c, d)


def band_autotune(G, NT, NB):
    G = np.ascontiguousarray(G, dtype=np.int32)
    return lib().orc_band_autotune(G.ctypes.data, NT, NB)


def rank_stat(G, NT, band):
    G = np.ascontiguousarray(G, dtype=np.int32)
    mn, mx, av = C.c_int(), C.c_int(), C.c_double()
    lib().orc_rank_stat(G.ctypes.data, NT, band, C.byref(mn), C.byref(mx), C.byref(av))
    return mn.value, mx.value, av.value


def check_dpotrf2(L, A0):
    N = L.shape[0]
    L = np.asfortranarray(L)
    A0 = np.asfortranarray(A0)
    return lib().orc_check_dpotrf2(N, L.ctypes.data, A0.ctypes.data)


def check_diff(L, L0):
    N = L.shape[0]
    L = np.asfortranarray(L)
    L0 = np.asfortranarray(L0)
    return lib().orc_check_diff(N, L.ctypes.data, L0.ctypes.data)


def dense_potrf(A):
    A = np.array(A, order="F", copy=True)
    info = lib().orc_dense_potrf(A.shape[0], A.ctypes.data)
    return A, info


def dense_logdet(L):
    L = np.asfortranarray(L)
    return lib().orc_dense_logdet(L.shape[0], L.ctypes.data)


class Oracle:
    """CPU TLR matrix following the reference algorithm tile by tile."""

    def __init__(self, N, NB, kind=2, sigma2=1.0, beta=0.1, nugget=1e-2, seed=0):
        self.N, self.NB = N, NB
        self.h = lib().orc_create(N, NB, kind, sigma2, beta, nugget, seed)
        self.NT = lib().orc_NT(self.h)
        self.band = 1

    def __del__(self):
        try:
            if self.h:
                lib().orc_free(self.h)
                self.h = None
        except Exception:
            pass

    def generate(self, tol=1e-8, maxrank=150, band_size=1):
        self.band = band_size
        self.maxrank = maxrank
        return lib().orc_generate(self.h, tol, maxrank, band_size)

    def regenerate(self, band):
        self.band = band
        return lib().orc_band_regenerate(self.h, band)

    def set_band(self, band):
        self.band = band
        return lib().orc_set_band(self.h, band)

    def rank_gather(self):
        G = np.zeros(self.NT * self.NT, dtype=np.int32)
        lib().orc_rank_gather(self.h, G.ctypes.data)
        return G

    def potrf(self, acc=1e-8, rk=0, maxrank=150, kmax=-1):
        tile = np.zeros(self.NT * self.NT, dtype=np.uint64)
        st = np.zeros(4, dtype=np.uint64)
        sec = C.c_double()
        info = lib().orc_dpotrf(self.h, acc, rk, maxrank, kmax, tile.ctypes.data, st.ctypes.data, C.byref(sec))
        self.tileops, self.stats, self.seconds = tile.reshape(self.NT, self.NT), st, sec.value
        return info

    def uncompress(self):
        D = np.zeros((self.N, self.N), order="F")
        lib().orc_uncompress(self.h, D.ctypes.data)
        return D

    def dense_exact(self):
        D = np.zeros((self.N, self.N), order="F")
        lib().orc_dense_exact(self.h, D.ctypes.data)
        return D

    def logdet(self):
        return lib().orc_logdet(self.h)

    def rank(self, m, n):
        return lib().orc_get_rank(self.h, m, n)

    def get_tile(self, m, n):
        buf = np.zeros(self.NB * self.NB * 2)
        r = lib().orc_get_tile(self.h, m, n, buf.ctypes.data)
        if m - n < self.band:
            return buf[: self.NB * self.NB].reshape(self.NB, self.NB, order="F"), self.NB
        U = buf[: self.NB * r].reshape(self.NB, r, order="F")
        V = buf[self.NB * r: 2 * self.NB * r].reshape(self.NB, r, order="F")
        return (U, V), r

    def trsm(self, B, trans=0):
        """L^{-1} B (trans 0) or L^{-T} B (trans 1) with the factor held."""
        X = np.array(B, dtype=np.float64, order="F", copy=True)
        if X.ndim == 1:
            X = X.reshape(-1, 1, order="F")
        lib().orc_tlr_trsm(self.h, trans, X.shape[1], X.ctypes.data, X.shape[0])
        return X.reshape(np.shape(B)) if np.ndim(B) == 1 else X

    def set_tile(self, m, n, data, rank):
        data = np.ascontiguousarray(data, dtype=np.float64)
        return lib().orc_set_tile(self.h, m, n, data.ctypes.data, rank)
