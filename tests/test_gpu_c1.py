"""GPU parity at the reference's CTest size (hicma_parsec/CMakeLists.txt:62-65):
N=2700, NB=270, eps 1e-8, maxrank 130, 2D sqexp; 3-flow (band 1) and 2-flow
(auto band).  The oracle is the CPU restatement in oracle/."""
import numpy as np
import pytest

import hicma_tlr as H
import tlr_oracle as O

pytestmark = pytest.mark.gpu

N, NB, U, EPS = 2700, 270, 130, 1e-8


@pytest.fixture(scope="module")
def exact():
    o = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0)
    A = o.dense_exact()
    Ld, info = O.dense_potrf(A)
    assert info == 0
    return A, O.dense_logdet(Ld)


def test_generation_matches_oracle(ctx):
    g = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    assert g.generate(EPS, U, 1) == 0
    o = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0)
    assert o.generate(EPS, U, 1) == 0
    assert np.array_equal(g.points(), O.points(N, 2, 0))
    for m in range(1, g.NT):
        for n in range(m):
            (Ug, Vg), rg = g.get_tile(m, n)
            (Uo, Vo), ro = o.get_tile(m, n)
            assert abs(rg - ro) <= 2, (m, n, rg, ro)
            D = Uo @ Vo.T
            err = np.linalg.norm(Ug @ Vg.T - D) / np.linalg.norm(D)
            assert err < 3 * EPS, (m, n, err)
    Dg, _ = g.get_tile(3, 3)
    Do, _ = o.get_tile(3, 3)
    np.testing.assert_allclose(Dg, Do, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("auto", [False, True])
def test_factorization_check(ctx, exact, auto):
    A, ld_dense = exact
    g = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    assert g.generate(EPS, U, 1) == 0
    band = 1
    if auto:
        band = g.auto_tune()
        # the dense band kernels (trsm, syrk, gemm cases i-iv) are reached only when band > 1
        assert 1 < band < g.NT
        g.regenerate(band)
    Aapprox = g.uncompress()
    info = g.potrf(EPS, 0, U)
    assert info == 0
    L = g.uncompress()
    res = O.check_dpotrf2(L, A)
    print(f"band={band} ||A-LL^T||/||A|| = {res:.3e}")
    assert res <= 10 * EPS
    L0, info0 = O.dense_potrf(Aapprox)
    assert info0 == 0
    diff = O.check_diff(L, L0)
    print(f"check_diff = {diff:.3e}")
    assert diff <= 10 * EPS
    ld = g.logdet()
    rel = abs(ld - ld_dense) / abs(ld_dense)
    print(f"logdet {ld!r} dense {ld_dense!r} rel {rel:.2e}")
    assert rel <= 1e-8


def test_rank_check_history_and_counters(ctx):
    """parsec_rank_check (rank_check.jdf:22-61), the rank history behind
    parsec_rank_print (rank_print.jdf), the per-stream op counts and the
    critical-path times of one factorization."""
    g = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    assert g.generate(EPS, U, 1) == 0
    before = g.rank_map().copy()
    g.rank_history(True)
    assert g.potrf(EPS, 0, U, want_ops=True, critical_path=True) == 0
    after = g.rank_map().copy()
    assert g.rank_check() == 0
    R = g.rank_map()
    for m in range(g.NT):
        assert R[m, m] == -1                      # band entries set to -1
        for n in range(m):
            assert R[m, n] == after[m, n] >= 0
    # the history: initial = before, final = after, min <= both <= max
    h = g.get_rank_history()
    for m in range(g.NT):
        for n in range(m):
            ini, lo, hi, fin = h[m, n]
            assert ini == before[m, n] and fin == after[m, n]
            assert lo <= min(ini, fin) and hi >= max(ini, fin)
    assert (h[:, :, 1] < h[:, :, 2]).any()        # ranks changed somewhere during the factorization
    g.rank_print()
    st, ex = g.stats, g.exec_stats
    assert int(ex.sum()) == int(st[0] + st[1]) == int(g.tileops.sum())
    assert (ex > 0).sum() >= 2
    cpt = g.critical_path_time
    assert cpt[0] > 0 and cpt[1] > 0 and cpt[2] > 0 and np.all(cpt[3:] == 0)
    g.close()
