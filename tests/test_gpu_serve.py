"""The serving loop (hicma_server_*): Gaussian log-likelihood requests of
mixed size run concurrently on one device; each result must be the same
alone and in the mix (bit for bit: every kernel's result is independent of
grid sizes and of concurrent work), the same as a direct call on a
context of its own, and correct against the exact dense covariance.
Admission by HBM footprint: a budget that fits one request at a time
serialises them; a request above the budget is refused."""
import threading

import numpy as np
import pytest

import hicma_tlr as H
import tlr_oracle as O

pytestmark = pytest.mark.gpu

REQS = [  # (N, sigma2, beta, seed)
    (4096, 1.0, 0.1, 1), (8192, 0.7, 0.05, 2), (6144, 1.8, 0.2, 3), (12288, 1.2, 0.03, 4),
    (4608, 0.5, 0.15, 5), (10240, 1.5, 0.08, 6),
]


def submit_all(srv, reqs):
    return [srv.submit(N, s2, b, 1e-2, seed=sd, zseed=100 + sd) for (N, s2, b, sd) in reqs]


def test_alone_equals_mix_equals_direct():
    with H.Server(0, slots=4, hbm_budget_gb=120.0) as srv:
        alone = [srv.wait(t) for t in [submit_all(srv, [r])[0] for r in REQS]]
        mix_t = submit_all(srv, REQS)
        mix = [srv.wait(t) for t in mix_t]
    for a, m, (N, s2, b, sd) in zip(alone, mix, REQS):
        assert a["status"] == 0 and m["status"] == 0, (a, m)
        assert (a["logdet"], a["quad"], a["band"]) == (m["logdet"], m["quad"], m["band"]), (N, a, m)
        assert m["loglik"] == pytest.approx(-0.5 * (m["quad"] + m["logdet"] + N * np.log(2 * np.pi)), rel=1e-14)
    # the same computation on a context of its own (four streams, all CUs)
    with H.Context(0) as ctx:
        for a, (N, s2, b, sd) in zip(alone[:3], REQS[:3]):
            g = H.TLRMatrix(ctx, N, 512, 256, H.SQEXP_2D, s2, b, 1e-2, sd)
            assert g.generate_first(1e-8, 256) == 0
            G = g.rank_gather()
            G[G == H.RANK_PENDING] = H.SERVE_PENDING_RANK
            band = H.band_autotune_model(G, g.NT, g.NB, H.SERVE_DENSE_TFLOPS,
                                         H.SERVE_LR_TFLOPS)  # the server's band choice,
            band = min(band, max(1, H.SERVE_BAND_BYTES // (g.NT * g.NB * g.NB * 8)))  # within its band memory
            if band != 1:
                g.regenerate(band)
            assert g.generate_finish(1e-8, 256) == 0
            assert g.potrf(1e-8, 0, 256) == 0
            ld, quad, ll = g.loglik(H.normal(100 + sd, N))
            assert (ld, quad) == (a["logdet"], a["quad"]) and band == a["band"]
            g.close()
    # overlap: in the mix some requests ran while others were on the device
    spans = sorted((m["t_total"] - m["t_run"], m["t_total"]) for m in mix)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_result_against_exact_covariance():
    N, s2, b, sd = 4096, 1.3, 0.1, 11
    z = np.random.default_rng(5).standard_normal(N)
    with H.Server(0, slots=2, hbm_budget_gb=40.0) as srv:
        res = srv.wait(srv.submit(N, s2, b, 1e-2, seed=sd, z=z))
    assert res["status"] == 0
    A = O.Oracle(N, 512, 2, s2, b, 1e-2, sd).dense_exact()
    S = A + np.tril(A, -1).T
    sign, ld = np.linalg.slogdet(S)
    assert sign > 0 and res["logdet"] == pytest.approx(ld, rel=1e-8)
    assert res["quad"] == pytest.approx(float(z @ np.linalg.solve(S, z)), rel=1e-6)


def test_admission_by_footprint():
    # a budget just above one 8192-point request: requests run one at a time
    with H.Server(0, slots=4, hbm_budget_gb=1000.0) as probe:
        one = probe.wait(probe.submit(8192, 1.0, 0.1, 1e-2, seed=1, zseed=1))
    assert one["status"] == 0
    budget = 1.2 * one["bytes"] / 1e9
    with H.Server(0, slots=4, hbm_budget_gb=budget) as srv:
        ts = [srv.submit(8192, 1.0, 0.1, 1e-2, seed=1 + i, zseed=1 + i) for i in range(3)]
        big = srv.submit(65536, 1.0, 0.1, 1e-2, seed=9, zseed=9)
        res = [srv.wait(t) for t in ts]
        rb = srv.wait(big)
    assert all(r["status"] == 0 for r in res)
    spans = sorted((r["t_total"] - r["t_run"], r["t_total"]) for r in res)
    for i in range(len(spans) - 1):
        assert spans[i + 1][0] >= spans[i][1] - 1e-3      # never two at once
    assert rb["status"] == -4                             # larger than the whole budget


def test_many_waiters():
    """Waiting from several threads while submitting (the load generator's pattern)."""
    out = {}
    with H.Server(0, slots=4, hbm_budget_gb=100.0) as srv:
        ts = submit_all(srv, REQS[:4])

        def waiter(t):
            out[t] = srv.wait(t)
        th = [threading.Thread(target=waiter, args=(t,)) for t in ts]
        for x in th:
            x.start()
        for x in th:
            x.join()
    assert all(out[t]["status"] == 0 for t in ts)


@pytest.mark.parametrize("N,beta", [(12288, 0.03), (8192, 0.1)])
def test_staged_generation_equals_starsh_gen(ctx, N, beta):
    """hicma_gen_first + hicma_gen_finish leave every off-band tile as
    STARSH_gen does (the same passes on the same tiles): bit for bit at band
    1, and outside the band after a parsec_band_regenerate in between."""
    ref = H.TLRMatrix(ctx, N, 512, 256, H.SQEXP_2D, 1.0, beta, 1e-2, 3)
    assert ref.generate(1e-8, 256, 1) == 0
    R = ref.rank_map()
    for band in (1, 4):
        g = H.TLRMatrix(ctx, N, 512, 256, H.SQEXP_2D, 1.0, beta, 1e-2, 3)
        assert g.generate_first(1e-8, 256) == 0
        pending = int(np.sum(g.rank_map() == H.RANK_PENDING))
        if band > 1:
            g.regenerate(band)
        assert g.generate_finish(1e-8, 256) == 0
        Gm = g.rank_map()
        for m in range(g.NT):
            for n in range(m - band + 1):
                assert Gm[m, n] == R[m, n], (band, m, n)
                if R[m, n] > 0:
                    (U1, V1), _ = g.get_tile(m, n)
                    (U0, V0), _ = ref.get_tile(m, n)
                    assert np.array_equal(U1, U0) and np.array_equal(V1, V0), (band, m, n)
        print(f"N={N} beta={beta} band {band}: {pending} tiles pending after the first sketch")
        g.close()
    ref.close()
