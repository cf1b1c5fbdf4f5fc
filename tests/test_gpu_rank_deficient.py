"""Low-rank x low-rank updates whose stacked right factor is exactly rank
deficient, built tile by tile through hicma_set_tile.

Three tile rows (NT = 3, band 1): the only low-rank gemm is
C(2,1) -= A(2,0) B(1,0)^T (hicma_parsec/HiCMA_dpotrf_L_2flow.jdf:929-1097).
With r_A > r_B, Qn = -Bu (Av^T Bv)^T has rank <= r_B < r_A, so Q = [Cv Qn] is
singular; column 0 of Av is orthogonal to Bv, which makes the FIRST column of
Qn (numerically) zero: a dependent column ahead of independent ones.  The
diagonal tile (0,0) is 4 I, so the trsm only halves V and the orthogonality
survives it.  Each shape reaches a different recompression kernel (stack
width s = r_C + r_A).  The GPU factor is compared with numpy's dense
Cholesky of the same matrix and with the oracle's factorization."""
import numpy as np
import pytest

import hicma_tlr as H
import tlr_oracle as O

pytestmark = pytest.mark.gpu
EPS = 1e-8


def orth(rng, n, k):
    q, _ = np.linalg.qr(rng.standard_normal((n, k)))
    return q


def build(NB, rc, ra, rb, seed):
    rng = np.random.default_rng(seed)
    diag = [4.0 * np.eye(NB)]
    for _ in range(2):
        S = rng.standard_normal((NB, NB)) * 0.01
        diag.append(4.0 * np.eye(NB) + S + S.T)
    # B = tile (1,0), A = tile (2,0), C = tile (2,1); U holds the singular values
    Bv = orth(rng, NB, rb)
    Bu = orth(rng, NB, rb) * np.linspace(0.5, 0.1, rb)
    M = rng.standard_normal((NB, ra))
    M[:, 0] -= Bv @ (Bv.T @ M[:, 0])           # Av[:, 0] orthogonal to span(Bv)
    Av = np.linalg.qr(M)[0]
    assert np.abs(Av[:, 0] @ Bv).max() < 1e-14
    Au = orth(rng, NB, ra) * np.linspace(0.5, 0.05, ra)
    Cv = orth(rng, NB, rc)
    Cu = orth(rng, NB, rc) * np.linspace(0.4, 0.02, rc)
    tiles = {(0, 0): diag[0], (1, 1): diag[1], (2, 2): diag[2], (1, 0): (Bu, Bv), (2, 0): (Au, Av), (2, 1): (Cu, Cv)}
    A = np.zeros((3 * NB, 3 * NB))
    for (m, n), t in tiles.items():
        blk = t if m == n else t[0] @ t[1].T
        A[m * NB:(m + 1) * NB, n * NB:(n + 1) * NB] = blk
        if m != n:
            A[n * NB:(n + 1) * NB, m * NB:(m + 1) * NB] = blk.T
    return tiles, A


def load(g, tiles):
    for (m, n), t in tiles.items():
        if m == n:
            g.set_tile(m, n, np.asfortranarray(t).ravel(order="F"), 0)
        else:
            g.set_tile(m, n, np.concatenate([t[0].ravel(order="F"), t[1].ravel(order="F")]), t[0].shape[1])


# (NB, rc, ra, rb): s = rc + ra selects k_rc_wave<32> (s <= 32), <48>, <64>
# (s <= 60), the four-wave k_rc_wide (60 < s <= 128) and k_lr_gemm (s > 128)
SHAPES = [(64, 3, 4, 2), (128, 20, 20, 6), (128, 30, 26, 8), (256, 40, 30, 10), (512, 80, 60, 16)]


@pytest.mark.parametrize("NB,rc,ra,rb", SHAPES)
def test_dependent_column_first(ctx, NB, rc, ra, rb):
    tiles, A = build(NB, rc, ra, rb, seed=NB + rc)
    N, cap = 3 * NB, NB // 2
    g = H.TLRMatrix(ctx, N, NB, cap, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    g.generate(EPS, cap, 1)          # allocates the band; every tile is overwritten below
    load(g, tiles)
    assert g.potrf(EPS, 0, cap) == 0
    assert g.jacobi_unconverged() == 0
    L = np.tril(g.uncompress())
    Ln = np.linalg.cholesky(A)
    err = np.linalg.norm(L - Ln) / np.linalg.norm(Ln)
    print(f"NB={NB} rc={rc} ra={ra} rb={rb}: ||L - chol(A)|| / ||chol(A)|| = {err:.2e}, rank {g.rank_map()[2, 1]}")
    assert err <= 1e-9
    # the new rank of C(2,1) is the number of singular values >= acc of the
    # updated tile (+-1 at the threshold)
    Cd = A[2 * NB:, NB:2 * NB] - Ln[2 * NB:, :NB] @ Ln[NB:2 * NB, :NB].T
    S = np.linalg.svd(Cd, compute_uv=False)
    assert abs(int(g.rank_map()[2, 1]) - int((S >= EPS).sum())) <= 1
    # the oracle on the same tiles
    o = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0)
    o.set_band(1)
    load(o, tiles)
    assert o.potrf(EPS, 0, cap) == 0
    Lo = np.tril(o.uncompress())
    assert np.linalg.norm(L - Lo) <= 1e-9 * np.linalg.norm(Lo)
    g.close()
