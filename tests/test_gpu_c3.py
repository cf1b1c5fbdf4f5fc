"""The batched-workload shape at a size whose dense checks are cheap:
statistics-3d-exp (C(r) = exp(-r/0.1) + 1e-2 nugget), tile 1728, maxrank 864
(= NB/2, the -D 4 setting of hicma_parsec/README.md:158-163), N = 13,824
(24^3 jittered grid, Morton order).  The auto-tuned band (5 of 8 tile rows)
leaves low-rank tiles of rank 133..770, so the recompressions stack up to
~1,500 columns in 1728-row tiles.  Checked with the device --check against
the exact matrix and the dense FP64 log-det."""
import pytest

import hicma_tlr as H

pytestmark = pytest.mark.gpu
N, NB, U, EPS = 13824, 1728, 864, 1e-8


def test_c3_shape_exp3d(ctx):
    A = H.TLRMatrix(ctx, N, NB, U, H.EXP_3D, 1.0, 0.1, 1e-2, 0)
    assert A.generate(EPS, U, 1) == 0
    G = A.rank_gather()
    rmin, rmax, _ = H.rank_stat(G, A.NT, 1)
    band = A.auto_tune()
    print(f"ranks {rmin}..{rmax}, band {band}", flush=True)
    assert 1 < band < A.NT
    A.regenerate(band)
    import time
    t0 = time.time()
    assert A.potrf(EPS, 0, U) == 0
    print(f"factorization {time.time() - t0:.1f}s", flush=True)
    assert A.jacobi_unconverged() == 0 and A.rank_overflows() == 0
    ld = A.logdet()
    L = H.Dense(ctx, N)
    X = H.Dense(ctx, N)
    L.uncompress(A)
    X.generate(A)
    bad, res = H.check_dpotrf2(ctx, L, X, EPS)
    print(f"||A - LL^T|| / ||A|| = {res:.3e}", flush=True)
    assert bad == 0 and res <= 10 * EPS
    assert X.potrf() == 0
    ldd = X.logdet()
    print(f"logdet {ld!r} dense {ldd!r}", flush=True)
    assert abs(ld - ldd) <= 1e-8 * abs(ldd)
    for D in (L, X, A):
        D.close()
