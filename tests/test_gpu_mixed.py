"""Mixed-precision storage of the factor (hicma_tlr_normf, hicma_tlr_demote):
N = 8,100 points of the 2D sqexp problem (beta 0.1, nugget 1e-2), tile 450,
maxrank 150, eps 1e-8.  The factor is computed in FP64, then off-band tiles
with ||L_mn||_F below thr32 are rounded to FP32 and those below thr16 to
BF16.  Checked: ||A||_F and the tile norms against numpy on the uncompressed
matrix; the tile classes against numpy norms of the FP64 tiles; every stored
element against numpy's rounding of the FP64 value (bit for bit: FP32 by
round to nearest even, BF16 through FP32 then nearest even on the upper 16
bits); the solve on the demoted factor against the oracle's FP64 solve on the
same widened values (1e-12); the log-likelihood against the FP64 factor's
(log-det bit-identical, z^T Sigma^-1 z within 1e-7 relative); and the
device --check residual ||A - L L^T|| / ||A|| <= 10 eps of the demoted factor
against the exact matrix."""
import numpy as np
import pytest

import hicma_tlr as H
import tlr_oracle as O
from roundings import bf16_round, f32_round

pytestmark = pytest.mark.gpu
N, NB, U, EPS = 8100, 450, 150, 1e-8


def make(ctx):
    return H.TLRMatrix(ctx, N, NB, U, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)


@pytest.fixture(scope="module")
def base(ctx):
    g = make(ctx)
    assert g.generate(EPS, U, 1) == 0
    normA, T = g.normf(tiles=True)
    Aap = g.uncompress()
    assert g.potrf(EPS, 0, U) == 0
    tiles = {}
    for m in range(1, g.NT):
        for n in range(m):
            tiles[(m, n)] = g.get_tile(m, n)
    z = np.random.default_rng(7).standard_normal(N)
    ld, qd, _ = g.loglik(z)
    yield dict(g=g, normA=normA, T=T, Aap=Aap, tiles=tiles, z=z, ld=ld, qd=qd)
    g.close()


def test_normf(base):
    A, T, NT = base["Aap"], base["T"], base["g"].NT
    d = np.diag(A)
    want = np.sqrt(2.0 * np.sum(np.tril(A, -1) ** 2) + np.sum(d * d))
    assert abs(base["normA"] - want) <= 1e-12 * want
    for n in range(NT):
        for m in range(n, NT):
            blk = A[m * NB:(m + 1) * NB, n * NB:(n + 1) * NB]
            if m == n:
                blk = np.tril(blk) + np.tril(blk, -1).T
            w = np.linalg.norm(blk)
            assert abs(T[m, n] - w) <= 1e-11 * w, (m, n, T[m, n], w)
        assert np.all(T[:n, n] == 0.0)


@pytest.mark.parametrize("t32,t16", [(EPS, 1e-3 * EPS), (1e-5, 1e-8)])
def test_demote(ctx, base, t32, t16):
    g, normA, tiles = base["g"], base["normA"], base["tiles"]
    thr32, thr16 = t32 * normA, t16 * normA
    h = make(ctx)
    try:
        h.copy_from(g)
        # classes from numpy norms of the FP64 tiles; rank-0 tiles stay FP64
        want_cls, exact = {}, True
        for key, ((Ut, Vt), r) in tiles.items():
            if r == 0:
                want_cls[key] = 0
                continue
            nrm = np.linalg.norm(Ut @ Vt.T)
            exact &= all(abs(nrm - t) > 1e-9 * t for t in (thr32, thr16))
            want_cls[key] = 2 if nrm < thr16 else (1 if nrm < thr32 else 0)
        counts = h.demote(thr32, thr16)
        want_counts = tuple(sum(1 for c in want_cls.values() if c == k) for k in range(3))
        print(f"thresholds {thr32:.3e} {thr16:.3e}: tiles FP64/FP32/BF16 {counts}, numpy {want_counts}")
        if exact:
            assert counts == want_counts
        if t32 > EPS:
            assert counts[1] > 0 and counts[2] > 0
        # every stored element is the rounding of the FP64 value
        for key, ((Ut, Vt), r) in tiles.items():
            (U2, V2), r2 = h.get_tile(*key)
            assert r2 == r
            cls = want_cls[key]
            rnd = (lambda x: x) if cls == 0 else (f32_round if cls == 1 else bf16_round)
            assert np.array_equal(U2, rnd(Ut)) and np.array_equal(V2, rnd(Vt)), (key, cls)
        # the solve reads the demoted tiles: same result as the oracle's FP64
        # solve on the widened values
        z = base["z"]
        ld2, qd2, _ = h.loglik(z)
        assert ld2 == base["ld"]
        rel = abs(qd2 - base["qd"]) / abs(base["qd"])
        print(f"z^T Sigma^-1 z: FP64 {base['qd']!r} mixed {qd2!r} relative difference {rel:.2e}")
        assert rel <= 1e-7
        o = O.Oracle(N, NB, 2, 1.0, 0.1, 1e-2, 0)
        o.set_band(h.band)
        for n in range(h.NT):
            for m in range(n, h.NT):
                d, r = h.get_tile(m, n)
                if m - n < h.band:
                    o.set_tile(m, n, np.asfortranarray(d).ravel(order="F"), 0)
                else:
                    o.set_tile(m, n, np.concatenate([d[0].ravel(order="F"), d[1].ravel(order="F")]), r)
        y = o.trsm(z.reshape(N, 1), 0)[:, 0]
        qo = float(y @ y)
        assert abs(qd2 - qo) <= 1e-12 * abs(qo)
        # the --check of the demoted factor against the exact matrix
        L, X = H.Dense(ctx, N), H.Dense(ctx, N)
        try:
            L.uncompress(h)
            X.generate(h)
            bad, res = H.check_dpotrf2(ctx, L, X, EPS)
        finally:
            L.close()
            X.close()
        print(f"||A - LL^T|| / ||A|| of the mixed factor = {res:.3e}")
        assert bad == 0 and res <= 10 * EPS
        # the factorization refuses demoted tiles; hicma_set_tile makes a tile FP64 again
        if counts[1] + counts[2]:
            with pytest.raises(RuntimeError):
                h.potrf(EPS, 0, U)
            key = next(k for k, c in want_cls.items() if c != 0)
            (Ut, Vt), r = tiles[key]
            h.set_tile(*key, np.concatenate([Ut.ravel(order="F"), Vt.ravel(order="F")]), r)
            (U3, V3), _ = h.get_tile(*key)
            assert np.array_equal(U3, Ut) and np.array_equal(V3, Vt)
        # STARSH_gen makes every tile FP64: the factorization runs again
        assert h.generate(EPS, U, 1) == 0
        assert h.potrf(EPS, 0, U) == 0
    finally:
        h.close()
