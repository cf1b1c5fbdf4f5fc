#!/usr/bin/env python3
"""Headline benchmark: TLR dpotrf seconds on one MI355X.

Workload: statistics-2d-sqexp, N=90,000 points (300x300 jittered grid, Morton
order, seed 0), NB=900 (100x100 tiles), eps 1e-8, maxrank 150, beta 0.1,
nugget 1e-2, band size auto-tuned -- the reference's testing_dpotrf_tlr
timed region (HiCMA_dpotrf_L only, hicma_parsec/testing_dpotrf.c:780-798).

A step is one HiCMA_dpotrf_L over the generated matrix, which is resident in
HBM: before every step the pristine compressed matrix is copied back
(outside the timed region), then the factorization is timed between device
synchronizations.

Also reported: the roofline of the dominant launch group (the low-rank gemm +
recompression kernels of each panel step, timed with HIP events on the
library's stream), and a CPU
baseline: the oracle (C restatement of the reference, MKL sequential per task,
OpenMP tasks on cores-1 threads) run on a bounded sample of the same matrix
(the first panel steps), extrapolated to the whole factorization by the
reference's op counts.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "hicma-parsec_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import hicma_tlr as H  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per recompression launch group at the default workload (N=90,000,
# NB=900, maxrank 150, eps 1e-8), measured with two rocprofv3 --pmc passes
# (FETCH_SIZE, WRITE_SIZE) summed over the group's kernels and scaled by the
# counters' reading of a device copy of known size (x2.0 and x1.0).
C2_TRAFFIC_BYTES_PER_GROUP = 6.399e9
# HBM the serving sweep's admission may reserve (of 288 GB): the rest covers
# what the stream-ordered pool holds beyond the reservations
SERVE_BUDGET_GB = 200.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may run on: the affinity mask, capped by a cgroup
    CPU quota when one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            n = min(n, max(1, -(-int(q) // int(per))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A, total_ops, budget_s):
    """The oracle (C restatement of the reference, MKL sequential per task,
    OpenMP tasks) on cores - 1 host threads, as the reference is run
    (-c = cores - 1): whole factorizations of the same compressed matrix,
    median of 3.  If one factorization would take longer than budget_s, the
    first panel steps are timed instead and scaled by op count."""
    import tlr_oracle as O

    cores = host_cores()
    threads = O.set_threads(max(1, cores - 1))

    def load():
        o = O.Oracle(A.N, A.NB, A.kind, 1.0, 0.1, 1e-2, 0)
        o.set_band(A.band)
        o.maxrank = A.maxrank
        for n in range(A.NT):
            for m in range(n, A.NT):
                d, r = A.get_tile(m, n)
                if m - n < A.band:
                    o.set_tile(m, n, np.asfortranarray(d).ravel(order="F"), 0)
                else:
                    U, V = d
                    o.set_tile(m, n, np.concatenate([U.ravel(order="F"), V.ravel(order="F")]), r)
        return o

    o = load()
    o.potrf(1e-8, 0, A.maxrank, kmax=1)
    t1, f1 = o.seconds, int(o.stats[0] + o.stats[1])
    est = t1 * total_ops / max(f1, 1)
    log(f"cpu: {threads} threads of {cores} cores; panel 0 took {t1:.2f}s; whole factorization ~{est:.1f}s")
    base = {"unit": "s", "cores": threads, "cores_available": cores, "cpu": cpu_model(),
            "blas": O.blas_version(), "kind": "port"}
    if est <= budget_s:
        runs = []
        for i in range(3):
            del o
            o = load()
            assert o.potrf(1e-8, 0, A.maxrank) == 0
            runs.append(o.seconds)
            log(f"cpu run {i}: {o.seconds:.3f}s")
        runs.sort()
        base.update(value=round(runs[1], 3), runs=[round(x, 3) for x in runs],
                    sample=("oracle (C restatement of the reference, MKL sequential per task, OpenMP tasks with "
                            "tile dependencies) factoring the same compressed matrix: whole HiCMA_dpotrf_L, "
                            "median of 3 runs"))
        return base
    kmax = int(max(1, min(A.NT, budget_s / max(t1, 1e-3))))
    del o
    o = load()
    o.potrf(1e-8, 0, A.maxrank, kmax=kmax)
    ts, fs = o.seconds, int(o.stats[0] + o.stats[1])
    base.update(value=round(ts * total_ops / fs, 3),
                sample=(f"oracle on panel steps 0..{kmax - 1} of the same compressed matrix: {ts:.2f}s for "
                        f"{fs:.3e} of {total_ops:.3e} reference op counts, scaled by op count"))
    return base


def serving_sweep(rates, nreq, slots, budget_gb, seed=2024):
    """The serving workload: Poisson arrivals of Gaussian log-likelihood
    requests (N = 512 k, k uniform in 8..128, sigma2 ~ U[0.5, 2],
    beta ~ U[0.02, 0.2], nugget 1e-2, tile 512, eps 1e-8, z ~ N(0, I) seeded),
    each swept rate with its own requests.  Reports completed requests per
    second (first arrival to last completion) and latency percentiles
    (submission to result)."""
    import threading
    rng = np.random.default_rng(seed)
    rows = []
    with H.Server(0, slots=slots, hbm_budget_gb=budget_gb) as srv:
        srv.wait(srv.submit(4096, 1.0, 0.1, 1e-2, seed=0, zseed=0))  # first-use costs out of the way
        log("serving: warm-up request done")
        rid = 1
        for rate in rates:
            reqs = []
            for _ in range(nreq):
                reqs.append((512 * int(rng.integers(8, 129)), float(rng.uniform(0.5, 2.0)),
                             float(rng.uniform(0.02, 0.2)), rid))
                rid += 1
            gaps = rng.exponential(1.0 / rate, size=nreq)
            res, sub, threads = [None] * nreq, [0.0] * nreq, []
            t0 = time.perf_counter()
            tnext = t0
            for i, (N, s2, b, sd) in enumerate(reqs):
                tnext += gaps[i]
                dt = tnext - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                sub[i] = time.perf_counter()
                tk = srv.submit(N, s2, b, 1e-2, seed=sd, zseed=sd)
                def waiter(i=i, tk=tk, N=N):
                    res[i] = srv.wait(tk)
                    log(f"  rate {rate}/s request {i} N={N}: status {res[i]['status']}, "
                        f"latency {res[i]['t_total']:.3f}s (queued {res[i]['t_queue']:.3f}s)")
                th = threading.Thread(target=waiter)
                th.start()
                threads.append(th)
            for th in threads:
                th.join()
            ok = [r for r in res if r["status"] == 0]
            done_at = max(sub[i] + res[i]["t_total"] for i in range(nreq))
            lat = np.array([r["t_total"] for r in ok])
            row = {"rate": rate, "requests": nreq, "failed": nreq - len(ok),
                   "req_per_s": round(len(ok) / (done_at - t0), 3),
                   "p50_s": round(float(np.percentile(lat, 50)), 4), "p99_s": round(float(np.percentile(lat, 99)), 4),
                   "mean_N": int(np.mean([r[0] for r in reqs])),
                   "mean_queue_s": round(float(np.mean([r["t_queue"] for r in ok])), 4)}
            log(f"serving rate {rate}/s: {row}")
            rows.append(row)
    return rows


def mixed_precision(A, P, eps, reps=5):
    """The factor A held in mixed precision (off-band tiles with ||L_mn||_F <
    eps ||A||_F in FP32, < 1e-3 eps ||A||_F in BF16; P holds the matrix before
    the factorization): tile counts, off-band bytes held, and the
    log-likelihood of one seeded z on the FP64 and on the mixed factor (time
    of the device solve, median of reps; difference of z^T Sigma^-1 z)."""
    normA = P.normf()
    z = np.random.default_rng(11).standard_normal(A.N)

    def timed():
        ts, r = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            r = A.loglik(z)
            ts.append(time.perf_counter() - t0)
        return r, float(np.median(ts))

    (ld64, q64, _), t64 = timed()
    _, T = A.normf(tiles=True)
    R = A.rank_map()
    thr32, thr16 = eps * normA, 1e-3 * eps * normA
    b64 = bmx = 0
    for m in range(A.NT):
        for n in range(m - A.band + 1):
            r = max(int(R[m, n]), 0)
            b = A.NB * 2 * r
            b64 += 8 * b
            bmx += b * (2 if T[m, n] < thr16 else (4 if T[m, n] < thr32 else 8))
    t0 = time.perf_counter()
    counts = A.demote(thr32, thr16)
    tdem = time.perf_counter() - t0
    (ld, q, _), tmx = timed()
    return {"normA_F": normA, "thresholds": [thr32, thr16], "tiles_f64_f32_bf16": list(counts),
            "offband_bytes_f64": b64, "offband_bytes_mixed": bmx, "demote_s": round(tdem, 4),
            "loglik_s_f64": round(t64, 5), "loglik_s_mixed": round(tmx, 5),
            "logdet_identical": bool(ld == ld64), "quad_rel_diff": abs(q - q64) / abs(q64)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=90000)
    ap.add_argument("--NB", type=int, default=900)
    ap.add_argument("--maxrank", type=int, default=150)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--schedule", type=int, default=None,
                    help="fixed schedule threshold (tasks per step below which the four-wave kernel runs)")
    ap.add_argument("--serve-rates", default="2,8,32,128")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-slots", type=int, default=4)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (the single-GPU share of the TLR Cholesky); "
              "--gpus > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    ctx = H.Context(0)
    A = H.TLRMatrix(ctx, a.N, a.NB, a.maxrank, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    t0 = time.time()
    info = A.generate(a.eps, a.maxrank, 1)
    tgen = time.time() - t0
    if info != 0:
        raise SystemExit(f"STARSH_gen info={info}")
    G = A.rank_gather()
    rmin, rmax, ravg = H.rank_stat(G, A.NT, 1)
    log(f"STARSH_gen {tgen:.2f}s ranks min/avg/max {rmin}/{ravg:.1f}/{rmax}")
    band = A.auto_tune()
    if band != 1:
        A.regenerate(band)
    log(f"band_size_auto_tuning -> {band}")
    pristine = H.TLRMatrix(ctx, a.N, a.NB, a.maxrank, H.SQEXP_2D, 1.0, 0.1, 1e-2, 0)
    pristine.copy_from(A)
    sched = 0
    if a.schedule is not None:
        sched = a.schedule
        ctx.set_wave_min_tasks(sched)
        log(f"schedule: four-wave kernel below {sched} tasks per step (given)")
    elif not a.no_autotune:
        t0 = time.time()
        sched, ts = A.autotune_schedule(a.eps, a.maxrank, reps=2)
        log(f"schedule autotuner: four-wave kernel below {sched} tasks per step ({ts:.3f}s; {time.time() - t0:.1f}s)")

    for w in range(a.warmup):
        A.copy_from(pristine)
        t0 = time.time()
        A.potrf(a.eps, 0, a.maxrank)
        log(f"warmup {w}: {time.time() - t0:.3f}s")
    times = []
    ctx.profile(True)
    for s in range(a.steps):
        A.copy_from(pristine)
        t0 = time.perf_counter()
        info = A.potrf(a.eps, 0, a.maxrank)
        times.append(time.perf_counter() - t0)
        log(f"step {s}: {times[-1]:.3f}s info={info}")
        if info != 0:
            raise SystemExit(f"HiCMA_dpotrf_L info={info}")
    kms, nl = ctx.profile_read()
    ctx.profile(False)
    rc_ops, rc_tasks, rc_bytes = A.recompress_stats()
    st = A.stats
    total_ops = int(st[0] + st[1])
    logdet = A.logdet()
    G2 = A.rank_gather()
    fmin, fmax, favg = H.rank_stat(G2, A.NT, band)
    sec = float(np.mean(times))
    launches_per_step = nl / a.steps
    kernel_s = kms * 1e-3 / a.steps
    achieved = rc_ops / kernel_s * 1e-12
    log(f"factor {sec:.3f}s; op count {total_ops:.3e} ({total_ops / sec * 1e-12:.2f} TFLOP/s); recompression "
        f"kernel {kernel_s:.3f}s over {launches_per_step:.0f} launches, {rc_tasks} tasks, {rc_ops:.3e} ops "
        f"-> {achieved:.3f} TFLOP/s; {rc_bytes / kernel_s * 1e-9:.1f} GB/s algorithmic")
    log(f"final ranks min/avg/max {fmin}/{favg:.1f}/{fmax}; logdet {logdet:.12e}; overflows {A.rank_overflows()}")
    assert np.isfinite(logdet) and A.rank_overflows() == 0

    mixed = mixed_precision(A, pristine, a.eps)
    log(f"mixed precision: tiles FP64/FP32/BF16 {mixed['tiles_f64_f32_bf16']}, off-band bytes "
        f"{mixed['offband_bytes_f64'] / 1e9:.2f} -> {mixed['offband_bytes_mixed'] / 1e9:.2f} GB, log-likelihood "
        f"{mixed['loglik_s_f64'] * 1e3:.2f} -> {mixed['loglik_s_mixed'] * 1e3:.2f} ms, z^T Sigma^-1 z relative "
        f"difference {mixed['quad_rel_diff']:.2e}")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(pristine, total_ops, a.cpu_seconds)
        log(f"cpu baseline {cpu['value']}s on {cpu['cores']} threads ({cpu['cpu']})")

    serving = None
    two_flow = int(A.two_flow)
    if not a.no_serve:
        # the server's slots get their own streams: free this context's first
        # (streams beyond the device's hardware queues share one and serialize)
        A.close()
        pristine.close()
        ctx.close()
        rates = [float(x) for x in a.serve_rates.split(",") if x]
        serving = {"slots": a.serve_slots, "hbm_budget_gb": SERVE_BUDGET_GB,
                   "workload": ("Poisson arrivals; N = 512 k, k ~ U{8..128}; sigma2 ~ U[0.5, 2], beta ~ U[0.02, 0.2], "
                                "nugget 1e-2; tile 512, eps 1e-8, maxrank 256; generate + factor + log-det + solve"),
                   "sweep": serving_sweep(rates, a.serve_requests, a.serve_slots, SERVE_BUDGET_GB)}

    out = {
        "metric": "TLR dpotrf seconds, 2D sqexp N=90k eps 1e-8",
        "value": round(sec, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"HiCMA_dpotrf_L, statistics-2d-sqexp N={a.N} (jittered grid, Morton order, seed 0), "
                         f"NB={a.NB}, eps={a.eps:g}, maxrank={a.maxrank}, beta=0.1, nugget=1e-2, "
                         f"band auto-tuned={band}"),
            "N": a.N, "NB": a.NB, "maxrank": a.maxrank, "band_size": band, "two_flow": two_flow,
            "ranks_initial": [rmin, round(ravg, 2), rmax], "ranks_final": [fmin, round(favg, 2), fmax],
            "op_count": total_ops, "logdet": logdet, "generation_s": round(tgen, 3),
            "schedule_wave_min_tasks": sched,
        },
        "roofline": {
            "bound": "mfma",
            "pipe": ("FP64 compute roofline: the group's QR panels, trailing updates and Jacobi rotations are "
                     "VALU FMAs, its tall products MFMA; gfx950's FP64 vector and matrix peaks are both 78.6 TFLOP/s"),
            "kernel": ("recompression launch group of a panel step: k_classify_wave + k_rc_wave<32|48|64> "
                       "(one wave per low-rank gemm task: QR + Jacobi-SVD) + k_rc_wide (stacks of 61..128 "
                       "columns, four waves per task), on three concurrent streams"),
            "achieved": round(achieved, 4),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
            "traffic": (C2_TRAFFIC_BYTES_PER_GROUP
                        if (a.N, a.NB, a.maxrank, a.eps) == (90000, 900, 150, 1e-8) else None),
            "traffic_unit": "bytes per launch group (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE)",
            "flops_per_launch": round(rc_ops / launches_per_step),
            "avg_launch_ms": round(kms / nl, 3) if nl else None,
            "algorithmic_GBs": round(rc_bytes / kernel_s * 1e-9, 2),
        },
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "serving": serving,
    }
    print(json.dumps(out))
    if a.no_serve:
        A.close()
        pristine.close()
    ctx.close()


if __name__ == "__main__":
    main()
