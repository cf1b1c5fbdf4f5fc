"""hicma_tlr_loglik_noalloc (the allocation-free entry point) against
hicma_tlr_loglik on the same factor and vector: two calls in a row, the
second on a smaller problem; results bit-identical (same kernels, same
inputs), and a problem above the scratch limit is refused."""
import ctypes as C

import numpy as np
import pytest

import hicma_tlr as H

pytestmark = pytest.mark.gpu
EPS = 1e-8


def both(ctx, g, z):
    L = H.lib()
    out = []
    for fn in (L.hicma_tlr_loglik, L.hicma_tlr_loglik_noalloc):
        X = H.DeviceArray(ctx, z.reshape(-1, 1))
        ld, qd, ll = C.c_double(), C.c_double(), C.c_double()
        assert fn(ctx.h, g.h, X.p, C.byref(ld), C.byref(qd), C.byref(ll)) == 0
        out.append((ld.value, qd.value, ll.value, X.get().ravel()))
        X.close()
    return out


def test_noalloc_matches_allocating_entry_point():
    with H.Context(0) as ctx:
        for N, NB, U in [(2700, 270, 130), (1600, 200, 100)]:     # the second call is smaller
            g = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
            assert g.generate(EPS, U, 1) == 0
            assert g.potrf(EPS, 0, U) == 0
            z = np.random.default_rng(N).standard_normal(N)
            (a_ld, a_q, a_ll, a_y), (b_ld, b_q, b_ll, b_y) = both(ctx, g, z)
            assert (a_ld, a_q, a_ll) == (b_ld, b_q, b_ll)
            np.testing.assert_array_equal(a_y, b_y)
            assert np.isfinite(b_ll) and b_q > 0
            g.close()
        # a problem whose workspace exceeds the 8 Mi-double scratch is refused
        # before anything runs: N = 200,000, NB = 2,000 needs 100 * 528 * 4096
        # doubles for the panels of L_kk^-1 alone
        big = H.TLRMatrix(ctx, 200000, 2000, 10, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
        X = H.DeviceArray(ctx, np.zeros((200000, 1)))
        ld = C.c_double()
        assert H.lib().hicma_tlr_loglik_noalloc(ctx.h, big.h, X.p, C.byref(ld), None, None) == -6
        X.close()
        big.close()
