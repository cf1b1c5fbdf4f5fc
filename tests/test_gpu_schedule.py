"""The schedule knob and its autotuner (hicma_set_wave_min_tasks,
hicma_autotune_schedule): every schedule gives a correct factor (the
reference's --check residual against the oracle's exact matrix), the same
one twice gives the same bits, schedules differ only by rounding, and the
autotuner leaves its input untouched and keeps a candidate it timed."""
import numpy as np
import pytest

import hicma_tlr as H
import tlr_oracle as O

pytestmark = pytest.mark.gpu
N, NB, U, EPS = 2700, 270, 130, 1e-8


def factor(ctx, P, thr):
    ctx.set_wave_min_tasks(thr)
    g = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    g.copy_from(P)
    assert g.potrf(EPS, 0, U) == 0
    out = (g.rank_map().copy(), np.tril(g.uncompress()), g.logdet())
    g.close()
    return out


def test_schedules_and_autotuner(ctx):
    P = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    assert P.generate(EPS, U, 1) == 0
    ranks0 = P.rank_map().copy()
    A = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0).dense_exact()
    runs = {}
    for thr in (0, 10 ** 6):                 # one-wave kernels / every task on four waves
        a, b = factor(ctx, P, thr), factor(ctx, P, thr)
        np.testing.assert_array_equal(a[1], b[1])
        assert a[2] == b[2]
        assert O.check_dpotrf2(a[1], A) <= 10 * EPS
        runs[thr] = a
    assert runs[0][2] == pytest.approx(runs[10 ** 6][2], rel=1e-12)
    assert np.abs(runs[0][1] - runs[10 ** 6][1]).max() <= 1e-9
    best, sec = P.autotune_schedule(EPS, U, reps=1)
    assert best in (0, 50, 100) and sec > 0
    np.testing.assert_array_equal(P.rank_map(), ranks0)   # input untouched
    got = factor(ctx, P, best)
    assert O.check_dpotrf2(got[1], A) <= 10 * EPS
    ctx.set_wave_min_tasks(0)
    P.close()
