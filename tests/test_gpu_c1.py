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
        if band != 1:
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
