"""The batched workloads' shapes at a size whose dense checks are cheap,
N = 13,824 (24^3 jittered grid, Morton order), tile 1728:
statistics-3d-exp (C(r) = exp(-r/0.1) + 1e-2 nugget), maxrank 864 (= NB/2,
the -D 4 setting of hicma_parsec/README.md:158-163): the auto-tuned band
(5 of 8 tile rows) leaves low-rank tiles of rank 133..770, so the
recompressions stack up to ~1,500 columns in 1728-row tiles;
statistics-3d-sqexp with beta 0.05, maxrank 500 (the -D 3 setting), whose
tiles on the first four sub-diagonals reach rank 480..806 at this tile size
(Morton order in 3-D keeps some distant blocks close), so they are generated
dense (initial band 5, the band_size argument of STARSH_gen) and
auto-tuning may only widen the band.  Checked with the
device --check against the exact matrix and the dense FP64 log-det."""
import pytest

import hicma_tlr as H

pytestmark = pytest.mark.gpu
N, NB, EPS = 13824, 1728, 1e-8


@pytest.mark.parametrize("kind,beta,U,band0", [(H.EXP_3D, 0.1, 864, 1), (H.SQEXP_3D, 0.05, 500, 5)])
def test_batched_shapes(ctx, kind, beta, U, band0):
    A = H.TLRMatrix(ctx, N, NB, U, kind, 1.0, beta, 1e-2, 0)
    assert A.generate(EPS, U, band0) == 0
    G = A.rank_gather()
    rmin, rmax, _ = H.rank_stat(G, A.NT, band0)
    band = max(band0, A.auto_tune())
    print(f"ranks {rmin}..{rmax}, band {band}", flush=True)
    assert 1 <= band < A.NT
    if band != band0:
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
