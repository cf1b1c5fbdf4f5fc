"""The device-side --check (STARSH_check into a dense device matrix, the
dense Cholesky, check_dpotrf2 / check_diff, norms and log-det;
hicma_parsec/testing_dpotrf.c:111-230, 640-903) against the oracle and numpy
on the same inputs, and the reference harness run end to end with --check."""
import os
import re
import subprocess

import numpy as np
import pytest

import hicma_tlr as H
import tlr_oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS = 1e-8


def lansy_np(X):
    L = np.tril(X)
    return np.linalg.norm(L + np.tril(L, -1).T)


@pytest.mark.parametrize("N,NB,U", [(2700, 270, 130), (2600, 270, 130)])
def test_device_checks_match_oracle(ctx, N, NB, U):
    g = H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    assert g.generate(EPS, U, 1) == 0
    band = g.auto_tune()
    assert 1 < band < g.NT
    g.regenerate(band)
    o = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0)
    A = o.dense_exact()
    # exact matrix generated on the device = the oracle's closed form
    D1 = H.Dense(ctx, N)
    D1.generate(g)
    np.testing.assert_allclose(D1.get(), A, rtol=1e-13, atol=1e-300)  # exp() rounding differs by a few ulps
    # STARSH_check = U V^T / band tiles assembled on the host from the tiles
    D2 = H.Dense(ctx, N)
    D2.uncompress(g)
    Aap = D2.get()
    ref = np.zeros((N, N))
    for n in range(g.NT):
        for m in range(n, g.NT):
            d, r = g.get_tile(m, n)
            rows, cols = min(NB, N - m * NB), min(NB, N - n * NB)
            blk = d[:rows, :cols] if m - n < band else (d[0] @ d[1].T)[:rows, :cols]
            ref[m * NB:m * NB + rows, n * NB:n * NB + cols] = blk
    ref = np.tril(ref)
    assert np.abs(Aap - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.all(np.triu(Aap, 1) == 0)
    # norms
    assert D1.lansy() == pytest.approx(lansy_np(A), rel=1e-12)
    assert D2.lansy(D1) == pytest.approx(lansy_np(Aap - A), rel=1e-10)
    # dense Cholesky and log-det = numpy
    Ln = np.linalg.cholesky(Aap + np.tril(Aap, -1).T)
    assert D2.potrf() == 0
    L0 = np.tril(D2.get())
    assert np.linalg.norm(L0 - Ln) <= 1e-12 * np.linalg.norm(Ln)
    assert D2.logdet() == pytest.approx(2 * np.log(np.diag(Ln)).sum(), rel=1e-13)
    # factor, then both checks = the oracle's checks on the same matrices
    assert g.potrf(EPS, 0, U) == 0
    D1.uncompress(g)
    L = D1.get()
    # threshold 10 eps (the reference harness passes eps itself,
    # hicma_parsec/testing_dpotrf.c:883, which a ragged last tile can exceed)
    bad, diff = H.check_diff(ctx, D1, D2, 10 * EPS)
    assert diff == pytest.approx(O.check_diff(L, L0), rel=1e-9)
    assert bad == 0 and diff <= 10 * EPS
    D2.generate(g)
    bad, res = H.check_dpotrf2(ctx, D1, D2, 10 * EPS)
    assert res == pytest.approx(O.check_dpotrf2(L, A), rel=1e-6)
    assert bad == 0 and res <= 10 * EPS
    # a planted error in L is seen by both
    Lp = L.copy()
    Lp[N - 5, 3] += 1e-4
    D1.set(Lp)
    bad, res_p = H.check_dpotrf2(ctx, D1, D2, EPS)
    assert bad == 1 and res_p == pytest.approx(O.check_dpotrf2(Lp, A), rel=1e-6)
    for X in (D1, D2, g):
        X.close()


def test_harness_check_run():
    """testing_dpotrf_tlr -N 2700 -t 270 -e 1e-8 -u 130 -D 2 -x: exit status 0
    (both checks pass) and the reference's R-LO record."""
    exe = os.path.join(ROOT, "hicma-parsec_amd", "testing_dpotrf_tlr")
    r = subprocess.run([exe, "-N", "2700", "-t", "270", "-e", "1e-8", "-u", "130", "-D", "2", "--nugget", "1e-2",
                        "-x", "--print-rank", "-c", "4"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0
    out = r.stdout
    assert "Factorization is CORRECT" in out and "SAME matrices" in out
    assert "Final Rank" in out and "Difference Rank, max - min" in out
    m = re.search(r"Relative error of approximation: (\S+)", out)
    assert m and float(m.group(1)) < 1e-7
    m = re.search(r"log det = (\S+) dense = (\S+) relative error = (\S+)", out)
    assert m and float(m.group(3)) <= 1e-8
    rlo = [ln for ln in out.splitlines() if ln.startswith("R-LO ")]
    assert len(rlo) == 1
    f = rlo[0].split()
    assert f[1:5] == ["2700", "2700", "270", "270"]
    # critical-path times (potrf, trsm, syrk) are measured, not zero
    tp = re.search(r"Total_critical_path_time (\S+), potrf (\S+), trsm (\S+), syrk (\S+)", out)
    assert tp and all(float(x) > 0 for x in tp.groups())


def test_harness_mixed_check_run():
    """testing_dpotrf_tlr --mixed -x (N 2,700, tile 270, band auto-tuned as in
    the run above): the factor's off-band tiles below eps ||A||_F are held in
    FP32/BF16 after the factorization (the far corner tile at this size) and
    the reference's checks pass on that factor (exit status 0)."""
    exe = os.path.join(ROOT, "hicma-parsec_amd", "testing_dpotrf_tlr")
    r = subprocess.run([exe, "-N", "2700", "-t", "270", "-e", "1e-8", "-u", "130", "-D", "2", "--nugget", "1e-2",
                        "-x", "--mixed"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0
    out = r.stdout
    assert "Factorization is CORRECT" in out and "SAME matrices" in out
    m = re.search(r"Mixed_precision\s+\|\|A\|\|_F= (\S+) thresholds (\S+) (\S+) tiles FP64= (\d+) FP32= (\d+) BF16= (\d+)",
                  out)
    assert m, "no Mixed_precision record"
    n64, n32, n16 = (int(x) for x in m.groups()[3:])
    assert n64 + n32 + n16 <= 10 * 9 // 2 and n32 + n16 >= 1
    m = re.search(r"log det = (\S+) dense = (\S+) relative error = (\S+)", out)
    assert m and float(m.group(3)) <= 1e-8
