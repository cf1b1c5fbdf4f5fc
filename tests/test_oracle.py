"""CPU tests of the oracle (oracle/), pinned by numpy fixtures, closed forms,
special cases and brute force on small inputs.  Reference files cited:
op_counts (hicma_parsec/op_counts.h:10-72), stats (hicma_parsec/hicma_internal.c:52-72),
auto-tuning (hicma_parsec/band_size_auto_tuning.jdf:201-318), checks
(hicma_parsec/testing_dpotrf.c:111-230)."""
import os

import numpy as np
import pytest

import tlr_oracle as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


@pytest.mark.parametrize("key,N,nd,seed", [("pts2_100", 100, 2, 0), ("pts3_64", 64, 3, 0),
                                           ("pts2_2700", 2700, 2, 0), ("pts3_1000_s7", 1000, 3, 7)])
def test_points_golden(key, N, nd, seed):
    np.testing.assert_array_equal(O.points(N, nd, seed), G[key])


def test_points_are_morton_ordered_jittered_grid():
    N, nd = 2700, 2
    P = O.points(N, nd, 0)
    s = 52
    c = np.floor(P * s).astype(np.int64)          # jitter < 0.5 spacing: cell = grid coordinate
    assert c.min() >= 0 and c.max() < s
    assert np.all(np.abs(P * s - c - 0.5) <= 0.4 + 1e-12)
    code = np.zeros(N, dtype=np.int64)
    for b in range(6):
        for d in range(nd):
            code |= ((c[d] >> b) & 1) << (b * nd + d)
    assert np.all(np.diff(code) > 0)               # strictly increasing Morton codes
    assert len(np.unique(code)) == N


def test_kernel_values_golden():
    P = G["pts2_2700"]
    out = np.zeros((20, 20), order="F")
    O.lib().orc_kernel_block(2, 20, 20, np.ascontiguousarray(P).ctypes.data, 2700, 270, 0, 1.0, 0.1, 1e-2,
                             out.ctypes.data, 20)
    np.testing.assert_allclose(out, G["k_sqexp2_blk"], rtol=1e-13, atol=0)
    O.lib().orc_kernel_block(2, 20, 20, np.ascontiguousarray(P).ctypes.data, 2700, 0, 0, 1.0, 0.1, 1e-2,
                             out.ctypes.data, 20)
    np.testing.assert_allclose(out, G["k_sqexp2_diag"], rtol=1e-13, atol=0)
    P3 = np.ascontiguousarray(G["pts3_1000_s7"])
    O.lib().orc_kernel_block(4, 20, 20, P3.ctypes.data, 1000, 100, 0, 1.0, 0.1, 1e-2, out.ctypes.data, 20)
    np.testing.assert_allclose(out, G["k_exp3_blk"], rtol=1e-13, atol=0)
    O.lib().orc_kernel_block(3, 20, 20, P3.ctypes.data, 1000, 100, 0, 2.0, 0.05, 1e-2, out.ctypes.data, 20)
    np.testing.assert_allclose(out, G["k_sqexp3_blk"], rtol=1e-13, atol=1e-300)


def test_op_counts_closed_forms():
    for a, b, c in [(900, 30, 40), (270, 130, 1), (1728, 864, 17)]:
        assert O.op_counts("m", a, b, c) == 2 * a * b * c
        assert O.op_counts("t", a, b, 1) == b * a * a
        assert O.op_counts("t", a, b, 2) == a * b * b
        assert O.op_counts("s", a) == 22 * a ** 3
        assert O.op_counts("r", a, b, 1) == a * a * b
        assert O.op_counts("r", a, b, 2) == a * b * b
        # Householder QR flop count 2mn^2 - 2n^3/3 + lower-order terms
        q = O.op_counts("q", a, b)
        lead = 2 * a * b * b - 2 * b ** 3 / 3
        assert abs(q - lead) <= 2 * a * b + 6 * b + 2 ** -20 * lead
        # Cholesky n^3/3
        cc = O.op_counts("c", a)
        assert abs(cc - (a ** 3 / 3 - a * a / 2 + a / 6)) <= 2
    # float32 rounding inside 'q' (op_counts.h:15): 2 n^3 / 3.0f is computed in float
    n, m = 301, 900
    f32 = int(np.float32(np.float32(2 * n ** 3) / np.float32(3.0)))
    f17 = int(np.float32(np.float32(17 * n) / np.float32(3.0)))
    assert O.op_counts("q", m, n) == 2 * m * n * n - f32 + 2 * m * n + f17


def test_rank_stat_closed_form():
    rng = np.random.default_rng(1)
    NT, band = 9, 2
    Gm = rng.integers(0, 100, size=(NT, NT)).astype(np.int32)   # [m, n]
    mask = np.array([[m - n >= band for n in range(NT)] for m in range(NT)])
    vals = Gm[mask]
    mn, mx, av = O.rank_stat(Gm.T.ravel(), NT, band)           # column major G[n*NT+m]
    assert (mn, mx) == (vals.min(), vals.max())
    assert av == pytest.approx(vals.mean(), rel=1e-15)


def test_autotune_special_cases():
    NT, NB = 12, 270
    zero = np.zeros(NT * NT, dtype=np.int32)
    assert O.band_autotune(zero, NT, NB) == 1          # TLR free -> keep only the diagonal dense
    full = np.full(NT * NT, NB // 2, dtype=np.int32)
    assert O.band_autotune(full, NT, NB) == NT         # TLR never cheaper -> dense
    # more expensive off-diagonal ranks never shrink the chosen band
    rng = np.random.default_rng(3)
    base = rng.integers(5, 40, size=NT * NT).astype(np.int32)
    b0 = O.band_autotune(base, NT, NB)
    assert O.band_autotune(base * 3, NT, NB) >= b0


def _c1(kind=2, N=2700, NB=270, maxrank=130, beta=0.1, nugget=1e-2):
    o = O.Oracle(N, NB, kind, 1.0, beta, nugget, 0)
    assert o.generate(1e-8, maxrank, 1) == 0
    return o


def test_compression_vs_svd():
    o = _c1()
    P = G["pts2_2700"]
    rel = G["c1_rank_rel"]
    for m in range(1, 10):
        for n in range(m):
            (U, V), r = o.get_tile(m, n)
            assert abs(r - rel[m, n]) <= 2, (m, n, r, rel[m, n])
            D = np.exp(-((P[:, m * 270:(m + 1) * 270, None] - P[:, None, n * 270:(n + 1) * 270]) ** 2).sum(0)
                       / 0.02)
            assert np.linalg.norm(D - U @ V.T) <= 3e-8 * np.linalg.norm(D)


@pytest.mark.parametrize("auto", [False, True])
def test_c1_check(auto):
    o = _c1()
    band = 1
    if auto:
        band = O.band_autotune(o.rank_gather(), o.NT, o.NB)
        assert band > 1
        o.regenerate(band)
    A = o.dense_exact()
    assert np.linalg.norm(A[np.tril_indices(2700, -1)]) ** 2 * 2 + np.linalg.norm(np.diag(A)) ** 2 == \
        pytest.approx(G["c1_fro"][0] ** 2, rel=1e-12)
    Aap = o.uncompress()
    assert o.potrf(1e-8, 0, 130) == 0
    L = o.uncompress()
    assert O.check_dpotrf2(L, A) <= 1e-7
    L0, info = O.dense_potrf(Aap)
    assert info == 0 and O.check_diff(L, L0) <= 1e-7
    ld = o.logdet()
    assert abs(ld - G["c1_logdet"][0]) <= 1e-8 * abs(G["c1_logdet"][0])
    st = o.stats
    assert st[0] + st[1] == st[2] + st[3] == o.tileops.sum()


def test_recompression_rank_brute_force():
    """NT = 3: the only low-rank gemm is gemm(2,1,0); its output rank equals the
    number of singular values >= acc of the dense update (+-1 at ties)."""
    N, NB = 192, 64
    o = O.Oracle(N, NB, 2, 1.0, 0.5, 1e-2, 0)
    assert o.generate(1e-8, 32, 1) == 0
    Aap = o.uncompress()
    Aap = Aap + np.tril(Aap, -1).T
    Ld = np.linalg.cholesky(Aap)
    Cd = Aap[128:192, 64:128] - Ld[128:192, 0:64] @ Ld[64:128, 0:64].T
    S = np.linalg.svd(Cd, compute_uv=False)
    expect = int((S >= 1e-8).sum())
    assert o.potrf(1e-8, 0, 32) == 0
    assert abs(o.rank(2, 1) - expect) <= 1
    L = o.uncompress()
    assert np.linalg.norm(np.tril(L) - Ld) <= 1e-8 * np.linalg.norm(Ld)


def test_nonspd_info():
    # a negative nugget makes the first pivot 1 - 2 < 0: info = 1 (k*NB + iinfo)
    o = O.Oracle(2700, 270, 2, 1.0, 0.1, -2.0, 0)
    assert o.generate(1e-8, 130, 1) == 0
    assert o.potrf(1e-8, 0, 130) == 1


def test_maxrank_too_small_is_info_minus_one():
    o = O.Oracle(2700, 270, 2, 1.0, 0.1, 1e-2, 0)
    assert o.generate(1e-8, 5, 1) == -1


def test_ragged_tiles():
    N, NB = 2600, 270
    o = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0)
    assert o.generate(1e-8, 130, 1) == 0
    A = o.dense_exact()
    assert o.potrf(1e-8, 0, 130) == 0
    assert O.check_dpotrf2(o.uncompress(), A) <= 1e-7


def test_smoke_config_compresses():
    # __graft_entry__.smoke() factors N=1600, NB=200, maxrank 100 on the GPU;
    # the oracle's generation must succeed there (N=1080, NB=108 does not)
    o = O.Oracle(1600, 200, 2, 1.0, 0.1, 1e-2, 0)
    assert o.generate(1e-8, 100, 1) == 0
    assert O.Oracle(1080, 108, 2, 1.0, 0.1, 1e-2, 0).generate(1e-8, 54, 1) == -1
