"""The factorization captured in a HIP graph and replayed: every replay gives
the eager result bit for bit (rank map, uncompressed factor, log-det), with
eager work of the same context (matrix restore, a solve) between replays, and
a new capture after the arguments change."""
import numpy as np
import pytest

import hicma_tlr as H

pytestmark = pytest.mark.gpu
N, NB, U, EPS = 2700, 270, 130, 1e-8


def state(g):
    return g.rank_map().copy(), g.uncompress(), g.logdet()


@pytest.mark.parametrize("auto", [False, True])
def test_graph_replays_equal_eager(ctx, auto):
    g = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    assert g.generate(EPS, U, 1) == 0
    if auto:
        band = g.auto_tune()
        assert band > 1
        g.regenerate(band)
    pristine = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    pristine.copy_from(g)
    assert g.potrf(EPS, 0, U) == 0
    ref = state(g)
    ref_ops = g.stats.copy()
    g.use_graph(True)
    z = np.random.default_rng(1).standard_normal(N)
    lik = None
    for rep in range(4):
        g.copy_from(pristine)
        assert g.potrf(EPS, 0, U) == 0
        got = state(g)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
        assert got[2] == ref[2], rep
        np.testing.assert_array_equal(g.stats, ref_ops)
        ll = g.loglik(z)                       # eager work between replays
        assert lik is None or ll == lik
        lik = ll
    # a changed argument (fixed-rank mode) is captured anew, then eager again
    g.copy_from(pristine)
    assert g.potrf(EPS, 40, U) == 0
    fixed = state(g)
    g.use_graph(False)
    g.copy_from(pristine)
    assert g.potrf(EPS, 40, U) == 0
    eager_fixed = state(g)
    np.testing.assert_array_equal(fixed[1], eager_fixed[1])
    for x in (g, pristine):
        x.close()
