// hicma_serve.hip -- continuous batching of Gaussian log-likelihood requests.
//
// No counterpart in the reference.  A request asks for log det Sigma and
// z^T Sigma^-1 z of a statistics-2d-sqexp covariance (its own N, kernel
// parameters, point seed and vector z): generation, band auto-tuning,
// TLR Cholesky and the forward solve, all on the device, the factor never
// leaving HBM.  The server runs `slots` workers, each a thread with its own
// single-stream context (one hardware queue) whose persistent grids and
// workspaces are sized for ncu / slots CUs, so several requests' kernels run
// on the device at once.  Requests wait in one FIFO queue and are admitted as
// soon as a worker is free and their HBM footprint fits the budget; a request
// whose auto-tuned band needs more memory than it reserved waits for it, and
// no new request is admitted meanwhile (running ones only release memory, so
// the wait ends).  Memory is allocated in stream order (hipMallocAsync), so a
// request finishing and freeing does not stall the others.
//
// Every kernel's result is independent of grid sizes and of what else runs:
// tasks are independent, reductions run in a fixed order.  A request's result
// is therefore the same alone and in any mix.
#include "internal.hpp"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

size_t hicma_solve_bytes(int N, int NB);  // hicma_solve.hip

// band choice of a request: the time model with this device's rates (dense
// tiles about 10 TFLOP/s as the batched band kernels run, recompressions
// about 3.5 TFLOP/s of reference op counts)
#define SERVE_DENSE_TFLOPS 60.0
#define SERVE_LR_TFLOPS 3.5
#define SERVE_BAND_BYTES ((size_t)4 << 30)  // dense band memory a request may take
#define SERVE_PENDING_RANK 160  // model rank of a tile the narrow sketch cannot vouch for (the second sketch's width)

namespace {

struct SvRequest {
    uint64_t id;
    hicma_lik_request r;
    std::vector<double> z;
    double t_submit, t_start;
    size_t bytes;     // reserved HBM
    bool exclusive;   // retried after running out of HBM: runs with nothing else
};

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct hicma_server {
    int device, nslots, cu_share;
    size_t budget, reserved;
    std::vector<hicma_ctx *> ctxs;  // one single-stream context per slot
    hicma_ctx *probe;  // ctxs[0]: its (read-only) parameters size the footprints
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<SvRequest *> queue;
    std::unordered_map<uint64_t, hicma_lik_result> done;
    uint64_t next_id;
    int growing;   // requests waiting for band memory
    int running;   // admitted, not finished
    bool exclusive_running;  // an exclusive request is running: admit nothing
    bool stop;
    std::vector<std::thread> workers;
};

// HBM a request needs with a dense band of `band` tile diagonals: the
// descriptor, the context workspace (generation / factorization / solve share
// one buffer that grows from one to the next: the old block is freed in
// stream order while the new one is allocated, so both are counted), the
// vector
static size_t footprint(hicma_server *S, const hicma_lik_request &r, int band)
{
    const int cap = r.maxrank;
    size_t ws = hicma_work_bytes(S->probe, r.N, r.NB, cap, band);
    size_t sv = hicma_solve_bytes(r.N, r.NB);
    return hicma_desc_bytes(r.N, r.NB, cap, band, 2) + ws + sv + sizeof(double) * (size_t)r.N + ((size_t)1 << 20);
}

static void publish(hicma_server *S, uint64_t id, const hicma_lik_result &res)
{
    S->done[id] = res;
    S->cv_done.notify_all();
}

// with S->mu held: the next request to run, or null
static SvRequest *admit(hicma_server *S)
{
    if (S->growing > 0 || S->exclusive_running) return nullptr;
    const double t = now_s();
    for (auto it = S->queue.begin(); it != S->queue.end();) {
        SvRequest *q = *it;
        if (q->exclusive) {  // waits at the head until nothing runs
            if (S->running > 0) return nullptr;
            S->queue.erase(it);
            S->reserved += q->bytes;
            S->running++;
            S->exclusive_running = true;
            q->t_start = t;
            return q;
        }
        if (q->bytes > S->budget) {  // can never fit
            hicma_lik_result res{};
            res.status = HICMA_ERR_NOMEM;
            res.bytes = q->bytes;
            publish(S, q->id, res);
            delete q;
            it = S->queue.erase(it);
            continue;
        }
        if (S->reserved + q->bytes <= S->budget) {
            S->queue.erase(it);
            S->reserved += q->bytes;
            S->running++;
            q->t_start = t;
            return q;
        }
        // the head has waited long enough: nothing may pass it
        if (it == S->queue.begin() && t - q->t_submit > 1.0) return nullptr;
        ++it;
    }
    return nullptr;
}

// with S->mu not held: reserve the footprint of a dense band of `band` tile
// diagonals for a running request admitted with a smaller one, and return the
// band granted.  It waits until running requests have released enough; if
// every running request is waiting here (none would release anything), it
// takes the widest band whose footprint fits what is left of the budget
// (band 1, the admitted footprint, always does), so the reservations never
// exceed the budget.
static int grow(hicma_server *S, SvRequest *q, int band)
{
    std::unique_lock<std::mutex> lk(S->mu);
    S->growing++;
    S->cv_work.notify_all();
    auto extra = [&](int b) {
        const size_t need = footprint(S, q->r, b);
        return need > q->bytes ? need - q->bytes : (size_t)0;
    };
    S->cv_work.wait(lk, [&] { return S->reserved + extra(band) <= S->budget || S->growing == S->running; });
    while (band > 1 && S->reserved + extra(band) > S->budget) band--;
    const size_t x = extra(band);
    S->reserved += x;
    q->bytes += x;
    S->growing--;
    S->cv_work.notify_all();
    return band;
}

static hicma_lik_result run(hicma_server *S, hicma_ctx *ctx, SvRequest *q)
{
    const hicma_lik_request &r = q->r;
    hicma_lik_result res{};
    res.band = 1;
    hicma_tlr_desc *A = nullptr;
    double *dz = nullptr;
    int rc = 0, info = 0, band = 1;
    rc = hicma_desc_create(ctx, &A, r.N, r.NB, r.maxrank, HICMA_KIND_SQEXP_2D, r.sigma2, r.beta, r.nugget, r.seed);
    // generation in two stages: the band is chosen from the narrow sketch's
    // ranks (a tile it cannot vouch for counted at SERVE_PENDING_RANK), so
    // the wider sketches run only on the tiles that stay low rank
    if (!rc) rc = hicma_gen_first(ctx, A, r.eps, r.maxrank, &info);
    if (!rc && info) rc = info;
    if (!rc) {
        const int NT = (r.N + r.NB - 1) / r.NB;
        std::vector<int> G((size_t)NT * NT);
        rc = parsec_rank_gather(ctx, A, G.data(), 1);
        for (int &g : G)
            if (g == HICMA_RANK_PENDING) g = SERVE_PENDING_RANK;
        if (!rc) band = hicma_band_autotune_model(G.data(), NT, r.NB, SERVE_DENSE_TFLOPS, SERVE_LR_TFLOPS);
        if (band < 1) rc = band;
        // at most SERVE_BAND_BYTES of dense band per request
        const int bmax = (int)std::max<size_t>(1, SERVE_BAND_BYTES / ((size_t)NT * r.NB * r.NB * sizeof(double)));
        if (band > bmax) band = bmax;
    }
    if (!rc && band > 1) band = grow(S, q, band);
    if (!rc && band > 1) rc = parsec_band_regenerate(ctx, HICMA_LOWER, A, NULL, NULL, NULL, band);
    if (!rc) rc = hicma_gen_finish(ctx, A, r.eps, r.maxrank, &info);
    if (!rc && info) rc = info;
    res.band = band;
    if (!rc) {
        int la = -1, tf = 0;
        info = HiCMA_dpotrf_L(ctx, HICMA_LOWER, A, A, A, r.eps, 0, r.maxrank, &la, band, 300, r.maxrank, 0, &tf,
                              NULL, NULL, NULL, 0);
        rc = info;
    }
    if (!rc) rc = hicma_dev_malloc(ctx, sizeof(double) * (size_t)r.N, (void **)&dz);
    if (!rc) rc = hicma_memcpy(ctx, dz, q->z.data(), sizeof(double) * (size_t)r.N, 1);
    if (!rc) rc = hicma_tlr_loglik(ctx, A, dz, &res.logdet, &res.quad, &res.loglik);
    if (dz) hicma_dev_free(ctx, dz);
    if (A) hicma_desc_destroy(A);
    hicma_ctx_release_ws(ctx);
    res.status = rc;
    res.bytes = q->bytes;
    const double t = now_s();
    res.t_queue = q->t_start - q->t_submit;
    res.t_run = t - q->t_start;
    res.t_total = t - q->t_submit;
    return res;
}

static void worker(hicma_server *S, hicma_ctx *ctx)
{
    for (;;) {
        SvRequest *q = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return (q = admit(S)) != nullptr || (S->stop && S->queue.empty()); });
            if (!q) break;
        }
        if (q->exclusive) {
            // nothing else runs: return the pool's cached blocks to the device
            hipMemPool_t mp;
            if (hipDeviceGetDefaultMemPool(&mp, ctx->device) == hipSuccess) (void)hipMemPoolTrimTo(mp, 0);
        }
        const hicma_lik_result res = run(S, ctx, q);
        // the request's memory is freed in stream order: wait for the frees
        // before its reservation goes to another request
        (void)hipStreamSynchronize(ctx->stream);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->reserved -= q->bytes;
            S->running--;
            if (q->exclusive) S->exclusive_running = false;
            if (res.status == HICMA_ERR_NOMEM && !q->exclusive) {
                // the device ran out of HBM that the reservations did not
                // account for (blocks the stream-ordered pool keeps): run it
                // again, alone, first in line
                (void)hipGetLastError();
                q->exclusive = true;
                q->bytes = footprint(S, q->r, 1);
                S->queue.push_front(q);
                q = nullptr;
            } else {
                publish(S, q->id, res);
            }
            S->cv_work.notify_all();
        }
        delete q;
    }
}

extern "C" int hicma_server_create(hicma_server **out, int device, int slots, double hbm_budget_gb)
{
    if (!out || slots < 1 || slots > 32 || !(hbm_budget_gb > 0)) return HICMA_ERR_ARG;
    int ncu = 0;
    if (hipSetDevice(device) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu < 1)
        return HICMA_ERR_HIP;
    hicma_server *S = new hicma_server();
    S->device = device;
    S->nslots = slots;
    S->budget = (size_t)(hbm_budget_gb * 1e9);
    S->next_id = 1;
    S->cu_share = ncu / slots > 0 ? ncu / slots : 1;
    // every context is created here, so the process holds exactly one stream
    // per slot (the device offers few hardware queues: streams beyond them
    // share one and serialize)
    for (int i = 0; i < slots; i++) {
        hicma_ctx *c = nullptr;
        const int rc = hicma_init_ex(&c, device, HICMA_CTX_SINGLE_STREAM | HICMA_CTX_STREAM_ORDERED_ALLOC,
                                     S->cu_share);
        if (rc) {
            hicma_server_destroy(S);
            return rc;
        }
        S->ctxs.push_back(c);
    }
    S->probe = S->ctxs[0];
    for (int i = 0; i < slots; i++) S->workers.emplace_back(worker, S, S->ctxs[i]);
    *out = S;
    return 0;
}

extern "C" int hicma_server_submit(hicma_server *S, const hicma_lik_request *r, uint64_t *ticket)
{
    if (!S || !r || !ticket || r->N <= 0 || r->NB <= 0) return HICMA_ERR_ARG;
    SvRequest *q = new SvRequest();
    q->r = *r;
    if (q->r.maxrank <= 0) q->r.maxrank = r->NB / 2;
    if (q->r.maxrank > r->NB / 2 || !(q->r.eps > 0)) {
        delete q;
        return HICMA_ERR_ARG;
    }
    q->z.resize((size_t)r->N);
    if (r->z) memcpy(q->z.data(), r->z, sizeof(double) * (size_t)r->N);
    else hicma_normal(r->zseed, r->N, q->z.data());
    q->r.z = nullptr;
    q->bytes = footprint(S, q->r, 1);
    q->t_submit = now_s();
    std::lock_guard<std::mutex> lk(S->mu);
    q->id = S->next_id++;
    *ticket = q->id;
    S->queue.push_back(q);
    S->cv_work.notify_all();
    return 0;
}

extern "C" int hicma_server_wait(hicma_server *S, uint64_t ticket, hicma_lik_result *res)
{
    if (!S || !res) return HICMA_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    if (ticket == 0 || ticket >= S->next_id) return HICMA_ERR_ARG;
    S->cv_done.wait(lk, [&] { return S->done.count(ticket) > 0; });
    *res = S->done[ticket];
    S->done.erase(ticket);
    return 0;
}

extern "C" int hicma_server_destroy(hicma_server *S)
{
    if (!S) return 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    for (auto &t : S->workers) t.join();
    for (hicma_ctx *c : S->ctxs) hicma_fini(c);
    for (SvRequest *q : S->queue) delete q;
    delete S;
    return 0;
}
