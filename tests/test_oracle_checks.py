"""Pins of the oracle's checkers against independent computations.

* ``check_dpotrf2`` / ``check_diff`` (hicma_parsec/testing_dpotrf.c:111-230)
  against numpy: the Frobenius norm of dplasma_dlansy on the lower triangle
  of a symmetric matrix counts each strictly-lower entry twice.  A checker
  that only looked at the diagonal, or returned ~0, fails the perturbation
  cases below.
* ``op_counts('o')`` (hicma_parsec/op_counts.h:10-72): side 1 against a brute
  count of the multiplies and adds of an unblocked Householder application
  (LAPACK dorm2r/dlarf), side 2 against the DORMQR (side = R) multiplication
  and addition counts published in LAPACK Working Note 41.
* ``rank_check`` (hicma_parsec/rank_check.jdf:22-61) against a direct count.
"""
import numpy as np
import pytest

import tlr_oracle as O


def lansy_lower(X):
    """dlansy('F', 'L') of the symmetric matrix whose lower triangle is X."""
    L = np.tril(X)
    S = L + np.tril(L, -1).T
    return np.linalg.norm(S)


def spd(n, rng):
    B = rng.standard_normal((n, n))
    return B @ B.T + n * np.eye(n)


@pytest.mark.parametrize("n", [7, 40, 129])
def test_check_dpotrf2_equals_numpy(n):
    rng = np.random.default_rng(n)
    A = spd(n, rng)
    L = np.linalg.cholesky(A)
    A0 = np.tril(A)                      # the checker reads the lower triangle only
    exact = O.check_dpotrf2(L, A0)
    assert exact < 1e-14
    # perturb one strictly-lower entry of L: residual = lansy(L'L'^T - A)/lansy(A)
    for (i, j) in [(n - 1, 0), (n // 2, n // 3)]:
        Lp = L.copy()
        Lp[i, j] += 1e-3 * abs(L[i, j]) + 1e-4
        want = lansy_lower(Lp @ Lp.T - A) / lansy_lower(A)
        got = O.check_dpotrf2(Lp, A0)
        assert got == pytest.approx(want, rel=1e-10)
        assert got > 1e-8
    # entries above the diagonal of L must be ignored
    Lu = L.copy()
    Lu[0, n - 1] = 5.0
    assert O.check_dpotrf2(Lu, A0) < 1e-14


@pytest.mark.parametrize("n", [9, 64])
def test_check_diff_equals_numpy(n):
    rng = np.random.default_rng(100 + n)
    L0 = np.tril(rng.standard_normal((n, n)))
    L = L0.copy()
    L[n - 1, 1] += 0.25                  # off-diagonal: counted twice
    L[2, 2] -= 0.5                       # diagonal: counted once
    want = np.sqrt(2 * 0.25 ** 2 + 0.5 ** 2) / lansy_lower(L)
    assert O.check_diff(L, L0) == pytest.approx(want, rel=1e-12)
    # upper entries of either argument are ignored
    Lu = L.copy()
    Lu[0, 3] = 7.0
    assert O.check_diff(Lu, L0) == pytest.approx(want, rel=1e-12)


def dorm2r_left_count(m, n, k):
    """Multiplies + adds of dorm2r(side L) applying k reflectors to an m x n
    matrix, counted statement by statement of dlarf: w = C^T v (dgemv),
    w *= tau, C -= v w^T (dger); reflector i has length m - i."""
    cnt = 0
    for i in range(k):
        mi = m - i
        cnt += 2 * mi * n          # dgemv: mi*n multiplies + mi*n adds
        cnt += n                   # tau * w
        cnt += 2 * mi * n          # dger: mi*n multiplies + mi*n adds
    return cnt


def lawn41_dormqr_right(m, n, k):
    """LAPACK Working Note 41, DORMQR with side = R: multiplications
    2nmk - mk^2 + mk + nk - k^2/2 + k/2, additions 2nmk - mk^2 + mk."""
    mul = 2 * n * m * k - m * k * k + m * k + n * k - k * k // 2 + k // 2
    add = 2 * n * m * k - m * k * k + m * k
    return mul + add


@pytest.mark.parametrize("m,n,k", [(900, 30, 45), (270, 130, 17), (1728, 500, 864), (64, 1, 1)])
def test_op_counts_ormqr(m, n, k):
    # side 1: C is m x n, Q from k reflectors of length <= m
    assert O.op_counts("o", m, n, k, 1) == dorm2r_left_count(m, n, k)
    # side 2 (the reference passes (rows, cols, k, 2) for the V side)
    assert O.op_counts("o", m, n, k, 2) == lawn41_dormqr_right(m, n, k)


def test_rank_check_counts_and_marks_band():
    rng = np.random.default_rng(4)
    NT = 11
    for band in (1, 2, 4):
        G = rng.integers(-3, 30, size=(NT, NT)).astype(np.int32)   # [m, n]
        want = sum(1 for m in range(band, NT) for n in range(0, m - band + 1) if G[m, n] < 0 and n < m)
        Gc = np.ascontiguousarray(G.T).ravel()                      # column major G[n*NT+m]
        got = O.lib().orc_rank_check(Gc.ctypes.data, NT, band)
        assert got == want
        out = Gc.reshape(NT, NT).T
        for m in range(NT):
            for n in range(NT):
                if n <= m and m - n < band:
                    assert out[m, n] == -1
                else:
                    assert out[m, n] == G[m, n]


@pytest.mark.parametrize("N,NB,auto", [(2700, 270, False), (2600, 270, True)])
def test_tlr_trsm_equals_dense_solves(N, NB, auto):
    """orc_tlr_trsm against scipy's dense triangular solves on the
    uncompressed factor, both directions, and the Gaussian log-likelihood
    pieces against the exact dense covariance."""
    from scipy.linalg import solve_triangular
    o = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0)
    assert o.generate(1e-8, 130, 1) == 0
    if auto:
        band = O.band_autotune(o.rank_gather(), o.NT, o.NB)
        assert band > 1
        o.regenerate(band)
    A = o.dense_exact()
    assert o.potrf(1e-8, 0, 130) == 0
    L = np.tril(o.uncompress())
    rng = np.random.default_rng(N)
    B = rng.standard_normal((N, 3))
    for trans in (0, 1):
        want = solve_triangular(L, B, lower=True, trans=trans)
        got = o.trsm(B, trans)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    # z^T Sigma^-1 z = ||L^-1 z||^2 and the log-det, against the exact matrix:
    # the compression error (1e-8 relative) is amplified by the condition
    # number of Sigma (nugget 1e-2) in the quadratic form, hence 1e-6
    z = B[:, 0]
    y = o.trsm(z, 0)
    quad = float(y @ y)
    S = A + np.tril(A, -1).T
    want_quad = float(z @ np.linalg.solve(S, z))
    assert quad == pytest.approx(want_quad, rel=1e-6)
    sign, want_ld = np.linalg.slogdet(S)
    assert sign > 0 and o.logdet() == pytest.approx(want_ld, rel=1e-8)


def test_storage_roundings_known_answers():
    """The checker of the mixed-precision tests against values worked by
    hand: BF16 keeps 8 significant bits (ulp 2^-7 at 1), ties go to the even
    significand, and a double is rounded to FP32 first, so a value just above
    a BF16 tie by less than half an FP32 ulp lands on the tie and goes to even."""
    from roundings import bf16_round, f32_round
    cases = [
        (1.0, 1.0),
        (1 + 2.0 ** -8, 1.0),                          # tie, even significand below
        (1 + 3 * 2.0 ** -8, 1 + 2.0 ** -6),            # tie, even significand above
        (1 + 2.0 ** -8 + 2.0 ** -20, 1 + 2.0 ** -7),   # above the tie
        (1 + 2.0 ** -8 + 2.0 ** -30, 1.0),             # FP32 first: lands on the tie
        (-2.5, -2.5),
        (np.pi, 3.140625),
        (2.0 ** -130, 2.0 ** -130),                    # FP32 subnormal, exact in BF16
        (0.0, 0.0),
    ]
    x = np.array([c[0] for c in cases])
    assert np.array_equal(bf16_round(x), np.array([c[1] for c in cases]))
    y = np.array([1 + 2.0 ** -24, 1 + 3 * 2.0 ** -24, 1 + 2.0 ** -23, 0.1])
    assert np.array_equal(f32_round(y), np.array([1.0, 1 + 2.0 ** -22, 1 + 2.0 ** -23, 0.10000000149011612]))
