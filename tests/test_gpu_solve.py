"""Triangular solves, Sigma^-1 B and the Gaussian log-likelihood on the device
against the oracle's solves (oracle/tlr_oracle.c orc_tlr_trsm, itself pinned
to scipy) on the same factor, and against dense solves with the uncompressed
factor.  Tolerances: 1e-12 relative for the solves (both sides use the same
factor; only the summation order differs)."""
import numpy as np
import pytest

import hicma_tlr as H
import tlr_oracle as O

pytestmark = pytest.mark.gpu
EPS = 1e-8


def factored(ctx, N, NB, U, kind=H.SQEXP_2D, beta=0.1, auto=False):
    g = H.TLRMatrix(ctx, N, NB, U, kind, 1.0, beta, 1e-2, 0)
    assert g.generate(EPS, U, 1) == 0
    if auto:
        band = g.auto_tune()
        if band != 1:
            g.regenerate(band)
    assert g.potrf(EPS, 0, U) == 0
    return g


def oracle_of(g):
    """The oracle holding the GPU's factor, tile by tile."""
    o = O.Oracle(g.N, g.NB, g.kind, 1.0, 0.1, 1e-2, 0)
    o.set_band(g.band)
    for n in range(g.NT):
        for m in range(n, g.NT):
            d, r = g.get_tile(m, n)
            if m - n < g.band:
                o.set_tile(m, n, np.asfortranarray(d).ravel(order="F"), 0)
            else:
                o.set_tile(m, n, np.concatenate([d[0].ravel(order="F"), d[1].ravel(order="F")]), r)
    return o


@pytest.mark.parametrize("N,NB,U,auto", [(2700, 270, 130, False), (2700, 270, 130, True), (2600, 270, 130, True),
                                         (1999, 200, 100, False)])
def test_solves_match_oracle(ctx, N, NB, U, auto):
    from scipy.linalg import solve_triangular
    g = factored(ctx, N, NB, U, auto=auto)
    o = oracle_of(g)
    rng = np.random.default_rng(N + NB)
    B = rng.standard_normal((N, 6))              # two passes of 4 right-hand sides
    L = np.tril(g.uncompress())
    for trans in (0, 1):
        got = g.trsm(B, trans)
        want = o.trsm(B, trans)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), trans
        dense = solve_triangular(L, B, lower=True, trans=trans)
        assert np.linalg.norm(got - dense) <= 1e-11 * np.linalg.norm(dense), trans
    X = g.potrs(B)
    want = o.trsm(o.trsm(B, 0), 1)
    assert np.linalg.norm(X - want) <= 1e-12 * np.linalg.norm(want)
    # Sigma X = B with Sigma = L L^T
    assert np.linalg.norm(L @ (L.T @ X) - B) <= 1e-10 * np.linalg.norm(B)
    # log-likelihood of one vector
    z = B[:, 0]
    ld, quad, ll = g.loglik(z)
    y = o.trsm(z, 0)
    assert quad == pytest.approx(float(y @ y), rel=1e-12)
    assert ld == g.logdet()
    assert ll == pytest.approx(-0.5 * (quad + ld + N * np.log(2 * np.pi)), rel=1e-14)
    # against the exact covariance (the compression error, amplified by cond(Sigma))
    A = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0).dense_exact()
    S = A + np.tril(A, -1).T
    assert quad == pytest.approx(float(z @ np.linalg.solve(S, z)), rel=1e-6)
    # deterministic
    assert g.loglik(z) == (ld, quad, ll)
    g.close()


def test_solve_large_tiles(ctx):
    """Tile 1728 with maxrank 864 (the batched shape's tile, the largest LDS
    footprint of the solve kernels), 2-D sqexp, band 1 (ranks 67..111)."""
    N, NB, U = 6912, 1728, 864
    g = factored(ctx, N, NB, U, kind=H.SQEXP_2D, beta=0.1)
    o = oracle_of(g)
    B = np.random.default_rng(7).standard_normal((N, 2))
    for trans in (0, 1):
        got, want = g.trsm(B, trans), o.trsm(B, trans)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    g.close()
