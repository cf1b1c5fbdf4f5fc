"""Full-size GPU properties at the benchmark workload (N=90,000 points, NB=900,
maxrank 150, eps 1e-8, statistics-2d-sqexp): the oracle cannot factor this
size in seconds, so the checks are size-independent properties of the result
-- bit-identical reruns (the schedule is concurrent, the result must not be),
a positive definite factor, and the log-determinant against the oracle's
(tests/golden/c2_logdet.json, written by tests/golden/make_c2_logdet.py)."""
import json
import os

import numpy as np
import pytest

import hicma_tlr as H

pytestmark = pytest.mark.gpu
N, NB, U, EPS = 90000, 900, 150, 1e-8


def test_c2_rerun_is_bit_identical(ctx):
    A = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    assert A.generate(EPS, U, 1) == 0
    B = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    B.copy_from(A)
    assert A.potrf(EPS, 0, U) == 0 and B.potrf(EPS, 0, U) == 0
    assert A.rank_overflows() == 0
    np.testing.assert_array_equal(A.rank_map(), B.rank_map())
    la, lb = A.logdet(), B.logdet()
    assert la == lb and np.isfinite(la)
    # SPD with eigenvalues in (1e-2, N): log det / N in (log 1e-2, log N)
    assert np.log(1e-2) * N < la < np.log(N) * N
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c2_logdet.json")))["logdet"]
    assert abs(la - ref) <= 1e-9 * abs(ref)
    for T in (A, B):
        T.close()
