"""CPU tests of the product library's host side: it loads, exports every entry
point declared in include/hicma_tlr.h, and its host-only routines (point set,
op counts, auto-tuning, rank statistics, argument errors) agree with the
golden fixtures and the oracle.  No GPU call is made here."""
import os
import re
import subprocess

import numpy as np
import pytest

import hicma_tlr as H
import tlr_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "hicma_tlr.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:unsigned long|int|void|double)\s+\**(\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = H.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(H.EXPORTS)


@pytest.mark.parametrize("key,N,nd,seed", [("pts2_100", 100, 2, 0), ("pts3_64", 64, 3, 0),
                                           ("pts2_2700", 2700, 2, 0), ("pts3_1000_s7", 1000, 3, 7)])
def test_points_golden(key, N, nd, seed):
    np.testing.assert_array_equal(H.points(N, nd, seed), G[key])


def test_op_counts_match_oracle():
    rng = np.random.default_rng(0)
    for op in "qctmsor":
        for _ in range(50):
            a, b, c = (int(x) for x in rng.integers(1, 2000, size=3))
            d = int(rng.integers(1, 3))
            if op in "tr":
                c = d
            assert H.op_counts(op, a, b, c, d) == O.op_counts(op, a, b, c, d), (op, a, b, c, d)


def test_band_autotune_matches_oracle():
    rng = np.random.default_rng(5)
    for NT, NB in [(10, 270), (30, 900), (16, 512)]:
        for scale in (1, 5, 20, 60):
            Gm = rng.integers(0, scale + 1, size=NT * NT).astype(np.int32)
            assert H.band_autotune(Gm, NT, NB) == O.band_autotune(Gm, NT, NB)
    assert H.band_autotune(np.zeros(100, np.int32), 10, 270) == 1
    assert H.band_autotune(np.full(100, 135, np.int32), 10, 270) == 10


def test_rank_stat_matches_oracle():
    rng = np.random.default_rng(2)
    Gm = rng.integers(0, 50, size=64).astype(np.int32)
    for band in (1, 2, 3):
        assert H.rank_stat(Gm, 8, band) == O.rank_stat(Gm, 8, band)


def test_init_without_gpu_is_an_error_not_a_crash():
    import ctypes as C
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    assert H.lib().hicma_init(C.byref(h), 0) < 0


def test_driver_usage():
    exe = os.path.join(ROOT, "hicma-parsec_amd", "testing_dpotrf_tlr")
    assert os.path.exists(exe)
    r = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--fixedacc" in r.stderr
    assert "--mixed" in r.stderr and "--check" in r.stderr and "--print-rank" in r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0


def test_weighted_band_autotune():
    """Weight 1 is the reference's rule; a heavier low-rank weight never
    narrows the band; a weight so large that any low-rank work loses makes
    every tile with nonzero rank dense (band NT)."""
    rng = np.random.default_rng(11)
    for NT, NB in [(10, 270), (30, 900), (64, 512)]:
        for scale in (1, 5, 20, 60, 160):
            # ranks decaying away from the diagonal, as covariance tiles do
            G = np.zeros((NT, NT), np.int32)
            for m in range(NT):
                for n in range(m):
                    G[n, m] = int(rng.integers(0, scale + 1) * np.exp(-(m - n) / 4.0))
            Gm = G.ravel()
            b1 = H.band_autotune(Gm, NT, NB)
            assert H.band_autotune(Gm, NT, NB, 1.0) == b1
            prev = b1
            for w in (2.0, 8.0, 64.0):
                bw = H.band_autotune(Gm, NT, NB, w)
                assert bw >= prev or bw == NT, (NT, NB, scale, w, bw, prev)
                prev = bw
    full = np.full(100, 20, np.int32)
    assert H.band_autotune(full, 10, 270, 1e9) == 10


def _model_time_bruteforce(G, NT, NB, b, rd, rl):
    """The time model of hicma_band_autotune_model evaluated task by task
    (O(NT^3) per band): the reference it is checked against."""
    rk = lambda m, n: max(int(G[n * NT + m]), 0)
    nb = float(NB)
    rd, rl = 1.0 / (rd * 1e12), 1.0 / (rl * 1e12)

    def case_v(ra, rb, rcf):
        s, rc = int(ra + rcf), int(rcf)
        return (2 * H.op_counts('q', NB, s) + H.op_counts('m', ra, rb, NB) + H.op_counts('m', ra, NB, rb) +
                H.op_counts('r', s, s, 2) + H.op_counts('s', s) + s * rc + H.op_counts('o', NB, rc, s, 1) +
                H.op_counts('o', rc, NB, s, 2))
    t = 5e-4 * NT if b > 1 else 0.0   # per panel step of the 2-flow form
    for k in range(NT):
        for m in range(k + 1, NT):
            r = rk(m, k)
            t += (3 * nb * nb * r if m - k >= b else 2 * nb ** 3) * rd
        for n in range(k + 1, NT):
            for m in range(n + 1, NT):
                ra, rb, rc = rk(m, k), rk(n, k), 1.25 * rk(m, n)   # C's rank grown during the factorization
                dA, dB, dC = m - k < b, n - k < b, m - n < b
                wide = ra + rc > 128
                if dA:
                    t += 2 * nb ** 3 * rd
                elif dC and dB:
                    t += 4 * nb * nb * ra * rd
                elif dC:
                    t += (4 * nb * ra * rb + 2 * nb * nb * rb) * rd
                elif dB:
                    t += case_v(ra, 0, rc) * (300 * rl if wide else rl) + 2 * nb * nb * ra * rd
                else:
                    t += case_v(ra, rb, rc) * (300 * rl if (wide or rb > 64) else rl)
    return t


def test_band_model_equals_bruteforce():
    """hicma_band_autotune_model's difference-array evaluation picks the
    band that a direct task-by-task evaluation of the same model picks."""
    rng = np.random.default_rng(3)
    for NT, NB in [(6, 270), (9, 512), (12, 900)]:
        for scale in (5, 60, 200):
            G = np.zeros((NT, NT), np.int32)
            for m in range(NT):
                for n in range(m):
                    G[n, m] = int(rng.integers(0, scale + 1) * np.exp(-(m - n) / 3.0))
            Gm = G.ravel()
            times = [_model_time_bruteforce(Gm, NT, NB, b, 10.0, 3.5) for b in range(1, NT + 1)]
            want = 1 + int(np.argmin(times))
            assert H.band_autotune_model(Gm, NT, NB, 10.0, 3.5) == want, (NT, NB, scale)
    assert H.band_autotune_model(np.zeros(100, np.int32), 10, 270) == 1
