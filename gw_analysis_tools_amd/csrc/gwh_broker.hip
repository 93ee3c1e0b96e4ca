/*
 * gwh_broker.hip -- the serving broker: continuous batching of likelihood and Fisher requests from many
 * host threads onto one GPU.
 *
 * The reference's samplers call the likelihood concurrently from sampler threads (OpenMP over chains,
 * mcmc_sampler.cpp:4585-4600; a std::thread pool, :122-330).  Here those callers submit requests
 * (parameters copied at submission) and wait on tickets; one dispatcher thread forms device batches:
 *
 *  - likelihood requests of the same model and precision, in arrival order, are concatenated into one
 *    multi-event batch (each instance keeps its own event) up to max_batch_instances;
 *  - a batch goes to whichever of two streams is free, with its own scratch and pinned staging, so a
 *    batch's upload and launch overlap the previous batch's kernels; requests arriving meanwhile are
 *    admitted into the next batch as soon as a stream frees (not only when the GPU is idle);
 *  - admission is by an HBM budget: resident events + the broker's scratch + the new batch's needs must
 *    fit, otherwise the batch waits for in-flight work to drain (and a request that cannot fit even on
 *    an idle device fails with GWH_ERR_NOMEM);
 *  - Fisher requests go to two worker threads, each with its own stream and scratch, concurrently with the
 *    likelihood batches; a worker merges the queued Fisher requests of the same event, model, basis,
 *    precision and step into one batch, and sizes its passes (points per pass) so that resident events +
 *    likelihood scratch + every Fisher scratch stay within the HBM budget (waiting for in-flight work when
 *    the budget is taken; refusing with GWH_ERR_NOMEM what cannot fit on an idle device);
 *  - every event a request reads is pinned from submission until its ticket is collected (gwh_broker_wait)
 *    or the broker is destroyed: gwh_event_destroy returns GWH_ERR_BUSY meanwhile.
 *
 * An instance's result (and a Fisher point's) does not depend on what it is batched with (per-event tiling,
 * fixed-order sums), so a request's results are bit-identical alone and in the mix.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gwh_engine.h"

namespace {

using clk = std::chrono::steady_clock;

enum { REQ_LOGLIKE = 0, REQ_FISHER = 1 };

struct Request {
    long long ticket = 0;
    int kind = REQ_LOGLIKE;
    int model = 0, precision = 0, n = 0;
    std::vector<gwh_params> params;
    std::vector<int> events; /* one per instance */
    double *out = nullptr;   /* host, n (likelihood) or n * dim * dim (Fisher) */
    double *sigma = nullptr; /* Fisher only, optional */
    int basis = 0, dim = 0;
    double eps = 0;
    clk::time_point submitted, finished;
    int status = 0;           /* 0 pending, else the GWH status */
    std::vector<int> pinned;  /* distinct events held for this request */
    std::condition_variable done; /* its waiter (under the broker lock): one wake-up per request */
};

constexpr int kFisherWorkers = 2;
constexpr int kFisherMaxPoints = 256; /* points of one merged Fisher batch */

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Scratch scratch;
    DevBuf d_params, d_events, d_logl;
    gwh_params *h_params = nullptr; /* pinned staging */
    int *h_events = nullptr;
    double *h_logl = nullptr;
    int cap = 0; /* instances the pinned staging holds */
    bool busy = false;
    std::vector<std::shared_ptr<Request>> batch;
    std::vector<int> host_events; /* the batch's instance -> event map (host copy for the work lists) */
};

struct Broker {
    size_t budget = 0;
    int max_n = 0;
    std::mutex m;
    std::condition_variable cv_work;
    std::deque<std::shared_ptr<Request>> queue;
    std::map<long long, std::shared_ptr<Request>> all;
    long long next_ticket = 1;
    bool stop = false;
    Lane lanes[2];
    std::thread th;
    /* Fisher requests: their own queue and workers, each with a stream and scratch */
    std::deque<std::shared_ptr<Request>> fisher_queue;
    std::condition_variable cv_fisher;
    FisherScratch fisher[kFisherWorkers];
    bool fisher_busy[kFisherWorkers] = {false, false};
    size_t fisher_reserved[kFisherWorkers] = {0, 0}; /* bytes a worker's admitted batch may grow its scratch to */
    std::thread fisher_th[kFisherWorkers];
    /* statistics */
    long long batches = 0, instances = 0, fishers = 0, max_merged = 0, admission_waits = 0;
    long long fisher_batches = 0, max_fisher_merged = 0;
    size_t peak_bytes = 0;
};

std::mutex g_brokers_mu;
std::map<int, Broker *> g_brokers;
int g_next_broker = 1;

size_t resident_bytes()
{
    Engine &E = gwh_engine();
    size_t s = 0;
    for (auto &ev : E.events)
        if (ev.live) s += ev.bytes;
    return s;
}

size_t lane_bytes(const Lane &L)
{
    const Scratch &S = L.scratch;
    return S.coef.cap + S.partial.cap + S.active.cap + S.work.cap + S.wcount.cap + L.d_params.cap + L.d_events.cap +
           L.d_logl.cap;
}

/* device bytes the broker holds: resident events, likelihood lanes and Fisher scratch (engine lock held) */
size_t used_bytes(const Broker &B)
{
    size_t u = resident_bytes() + lane_bytes(B.lanes[0]) + lane_bytes(B.lanes[1]);
    for (int w = 0; w < kFisherWorkers; w++) u += std::max(B.fisher[w].bytes(), B.fisher_reserved[w]);
    return u;
}

/* pin the distinct events of a request (engine lock held) */
int pin_events(Engine &E, Request &r, const std::vector<int> &evs)
{
    for (int e : evs)
        if (e < 0 || e >= GWH_MAX_EVENTS || !E.events[e].live)
            return gwh_set_error(GWH_ERR_INVALID, "unknown event in request");
    for (int e : evs) {
        bool seen = false;
        for (int q : r.pinned) seen = seen || (q == e);
        if (seen) continue;
        r.pinned.push_back(e);
        E.events[e].pins++;
    }
    return GWH_OK;
}

void unpin_events(Request &r)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    for (int e : r.pinned)
        if (E.events[e].pins > 0) E.events[e].pins--;
    r.pinned.clear();
}

/* device bytes lane L holds after a batch of n instances over `events` (the buffers launch_batch and the
 * batched launch ensure, grown as gwh_grown_capacity says; engine lock held) */
size_t lane_bytes_after(const Lane &L, const std::vector<int> &events, int n)
{
    Engine &E = gwh_engine();
    size_t maxtiles = 0, tiles = 0;
    for (int i = 0; i < n; i++) {
        const Event &ev = E.events[events[i]];
        size_t t = gwh_tiles_of(ev.desc);
        tiles += t;
        if (t > maxtiles) maxtiles = t;
    }
    const Scratch &S = L.scratch;
    const size_t m = (size_t)n;
    return gwh_grown_capacity(m * sizeof(InstCoef), S.coef.cap) +
           gwh_grown_capacity(m * maxtiles * GWH_PARTIAL_STRIDE * sizeof(double), S.partial.cap) +
           gwh_grown_capacity(m * sizeof(int), S.active.cap) + gwh_grown_capacity(tiles * sizeof(int2), S.work.cap) +
           gwh_grown_capacity(GWH_NCLASS * sizeof(int), S.wcount.cap) +
           gwh_grown_capacity(m * sizeof(gwh_params), L.d_params.cap) +
           gwh_grown_capacity(m * sizeof(int), L.d_events.cap) + gwh_grown_capacity(m * sizeof(double), L.d_logl.cap);
}

void finish(Broker &B, const std::shared_ptr<Request> &r, int status)
{
    r->finished = clk::now();
    std::lock_guard<std::mutex> lk(B.m);
    r->status = status;
    r->done.notify_all();
}

/* launch one likelihood batch on lane L (engine lock held by the caller) */
int launch_batch(Broker &B, Lane &L)
{
    Engine &E = gwh_engine();
    int n = 0;
    for (auto &r : L.batch) n += r->n;
    int rc;
    if (n > L.cap) {
        if (L.h_params) (void)hipHostFree(L.h_params);
        if (L.h_events) (void)hipHostFree(L.h_events);
        if (L.h_logl) (void)hipHostFree(L.h_logl);
        L.h_params = nullptr;
        L.h_events = nullptr;
        L.h_logl = nullptr;
        L.cap = 0;
        GWH_HIP(hipHostMalloc((void **)&L.h_params, (size_t)n * sizeof(gwh_params), hipHostMallocDefault));
        GWH_HIP(hipHostMalloc((void **)&L.h_events, (size_t)n * sizeof(int), hipHostMallocDefault));
        GWH_HIP(hipHostMalloc((void **)&L.h_logl, (size_t)n * sizeof(double), hipHostMallocDefault));
        L.cap = n;
    }
    L.host_events.resize(n);
    int k = 0;
    for (auto &r : L.batch) {
        memcpy(L.h_params + k, r->params.data(), (size_t)r->n * sizeof(gwh_params));
        for (int i = 0; i < r->n; i++) L.host_events[k + i] = r->events[i];
        k += r->n;
    }
    memcpy(L.h_events, L.host_events.data(), (size_t)n * sizeof(int));
    if ((rc = L.d_params.ensure((size_t)n * sizeof(gwh_params))) != GWH_OK) return rc;
    if ((rc = L.d_events.ensure((size_t)n * sizeof(int))) != GWH_OK) return rc;
    if ((rc = L.d_logl.ensure((size_t)n * sizeof(double))) != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(L.d_params.p, L.h_params, (size_t)n * sizeof(gwh_params), hipMemcpyHostToDevice, L.stream));
    GWH_HIP(hipMemcpyAsync(L.d_events.p, L.h_events, (size_t)n * sizeof(int), hipMemcpyHostToDevice, L.stream));
    const int model = L.batch[0]->model, prec = L.batch[0]->precision;
    rc = gwh_loglike_launch(L.d_events.as<const int>(), L.host_events[0], model, prec,
                            L.d_params.as<const gwh_params>(), n, L.d_logl.as<double>(), L.stream, false, nullptr,
                            L.host_events.data(), 0, &L.scratch);
    if (rc != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(L.h_logl, L.d_logl.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, L.stream));
    GWH_HIP(hipEventRecord(L.done, L.stream));
    size_t used = used_bytes(B);
    std::lock_guard<std::mutex> lk(B.m);
    B.batches++;
    B.instances += n;
    if ((long long)L.batch.size() > B.max_merged) B.max_merged = (long long)L.batch.size();
    if (used > B.peak_bytes) B.peak_bytes = used;
    return GWH_OK;
}

void reap(Broker &B, Lane &L)
{
    int k = 0;
    for (auto &r : L.batch) {
        memcpy(r->out, L.h_logl + k, (size_t)r->n * sizeof(double));
        k += r->n;
        finish(B, r, GWH_OK);
    }
    L.batch.clear();
    L.busy = false;
}

bool compatible_fisher(const Request &a, const Request &b)
{
    return a.events[0] == b.events[0] && a.model == b.model && a.basis == b.basis && a.precision == b.precision &&
           a.dim == b.dim && a.eps == b.eps;
}

void fisher_worker(Broker *bp, int w)
{
    Broker &B = *bp;
    Engine &E = gwh_engine();
    FisherScratch &F = B.fisher[w];
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(B.m);
            B.cv_fisher.wait(lk, [&] { return B.stop || !B.fisher_queue.empty(); });
            if (B.fisher_queue.empty()) return; /* stop requested and nothing left */
            /* the head and every compatible queued request, in arrival order, up to kFisherMaxPoints */
            auto head = B.fisher_queue.front();
            int n = 0;
            for (auto it = B.fisher_queue.begin(); it != B.fisher_queue.end();) {
                if (compatible_fisher(**it, *head) && (n == 0 || n + (*it)->n <= kFisherMaxPoints)) {
                    n += (*it)->n;
                    batch.push_back(*it);
                    it = B.fisher_queue.erase(it);
                } else {
                    ++it;
                }
            }
            B.fisher_busy[w] = true;
        }
        const Request &h = *batch[0];
        std::vector<gwh_params> params;
        bool want_sigma = false;
        for (auto &r : batch) {
            params.insert(params.end(), r->params.begin(), r->params.end());
            want_sigma = want_sigma || (r->sigma != nullptr);
        }
        const int n = (int)params.size(), dim = h.dim;
        /* admission: shrink the operand pass until this worker's scratch fits beside everything else */
        int st = GWH_OK;
        for (;;) {
            bool others_busy, fits;
            size_t mo = (size_t)1 << 30, need = 0;
            {
                /* size and reserve in one critical section: the dispatcher's admission, which holds the
                   engine lock while it launches, sees the reservation before this worker allocates */
                std::lock_guard<std::recursive_mutex> le(E.mu);
                std::lock_guard<std::mutex> lk(B.m);
                B.fisher_reserved[w] = 0;
                const size_t other = used_bytes(B) - F.bytes();
                const size_t avail = (B.budget > other) ? B.budget - other : 0;
                others_busy = B.lanes[0].busy || B.lanes[1].busy;
                for (int q = 0; q < kFisherWorkers; q++) others_busy = others_busy || (q != w && B.fisher_busy[q]);
                for (;;) {
                    need = gwh_fisher_batch_bytes(h.events[0], h.model, h.basis, h.precision, params.data(), n, mo,
                                                  F, want_sigma);
                    if (need == 0 || need <= avail || mo < 4096) break;
                    mo /= 2;
                }
                fits = (need == 0 || need <= avail);
                if (fits) B.fisher_reserved[w] = need;
            }
            if (!fits) {
                if (others_busy) { /* wait for in-flight work, then size again */
                    {
                        std::lock_guard<std::mutex> lk(B.m);
                        B.admission_waits++;
                    }
                    std::this_thread::sleep_for(std::chrono::microseconds(200));
                    continue;
                }
                if (F.bytes() > 0) { /* give back this worker's own buffers and size from nothing */
                    std::lock_guard<std::recursive_mutex> le(E.mu);
                    std::lock_guard<std::mutex> lk(B.m);
                    F.release();
                    continue;
                }
                st = gwh_set_error(GWH_ERR_NOMEM, "Fisher request exceeds the HBM budget");
                break;
            }
            F.max_operand = mo;
            break;
        }
        std::vector<double> out((size_t)n * dim * dim), sig(want_sigma ? (size_t)n * dim : 0);
        if (st == GWH_OK)
            st = gwh_fisher_batch_with(h.events[0], h.model, h.basis, h.precision, params.data(), n, dim, h.eps,
                                       out.data(), want_sigma ? sig.data() : nullptr, F);
        {
            std::lock_guard<std::recursive_mutex> le(E.mu);
            std::lock_guard<std::mutex> lk(B.m);
            B.fishers += (long long)batch.size();
            B.fisher_batches++;
            if ((long long)batch.size() > B.max_fisher_merged) B.max_fisher_merged = (long long)batch.size();
            const size_t used = used_bytes(B);
            if (used > B.peak_bytes) B.peak_bytes = used;
            B.fisher_reserved[w] = 0;
            B.fisher_busy[w] = false;
        }
        size_t k = 0;
        for (auto &r : batch) {
            if (st == GWH_OK) {
                memcpy(r->out, out.data() + k * dim * dim, (size_t)r->n * dim * dim * sizeof(double));
                if (r->sigma) memcpy(r->sigma, sig.data() + k * dim, (size_t)r->n * dim * sizeof(double));
            }
            k += r->n;
            finish(B, r, st);
        }
    }
}

void dispatcher(Broker *bp)
{
    Broker &B = *bp;
    Engine &E = gwh_engine();
    for (;;) {
        /* 1. reap finished lanes (no broker lock while querying the device) */
        for (Lane &L : B.lanes)
            if (L.busy && hipEventQuery(L.done) == hipSuccess) reap(B, L);
        Lane *free_lane = nullptr;
        for (Lane &L : B.lanes)
            if (!L.busy) {
                free_lane = &L;
                break;
            }
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(B.m);
            bool any_busy = B.lanes[0].busy || B.lanes[1].busy;
            if (B.queue.empty() || !free_lane) {
                if (B.stop && B.queue.empty() && !any_busy) return;
                if (!any_busy) {
                    B.cv_work.wait(lk, [&] { return B.stop || !B.queue.empty(); });
                    continue;
                }
                /* work in flight: poll shortly (the next submission or a completion ends the wait) */
                B.cv_work.wait_for(lk, std::chrono::microseconds(20));
                continue;
            }
            auto head = B.queue.front();
            /* continuous batching: compatible requests in arrival order up to max_n instances */
            int n = 0;
            for (auto it = B.queue.begin(); it != B.queue.end();) {
                auto &r = *it;
                if (r->model == head->model && r->precision == head->precision && (n == 0 || n + r->n <= B.max_n)) {
                    n += r->n;
                    batch.push_back(r);
                    it = B.queue.erase(it);
                    if (n >= B.max_n) break;
                } else {
                    ++it;
                }
            }
        }
        /* admission by the HBM budget */
        std::vector<int> evs;
        int n = 0;
        for (auto &r : batch) {
            evs.insert(evs.end(), r->events.begin(), r->events.end());
            n += r->n;
        }
        {
            std::lock_guard<std::recursive_mutex> lk(E.mu);
            size_t used = used_bytes(B);
            size_t mine = lane_bytes(*free_lane);
            if (used - mine + lane_bytes_after(*free_lane, evs, n) > B.budget) {
                bool other_busy = B.lanes[0].busy || B.lanes[1].busy;
                {
                    std::lock_guard<std::mutex> lq(B.m);
                    for (int q = 0; q < kFisherWorkers; q++) other_busy = other_busy || B.fisher_busy[q];
                }
                if (other_busy) {
                    /* wait for in-flight work, then retry with the batch back at the queue head */
                    std::lock_guard<std::mutex> lq(B.m);
                    for (auto it = batch.rbegin(); it != batch.rend(); ++it) B.queue.push_front(*it);
                    B.admission_waits++;
                    batch.clear();
                } else {
                    /* nothing in flight: idle Fisher scratch is given back before refusing */
                    {
                        std::lock_guard<std::mutex> lq(B.m);
                        for (int q = 0; q < kFisherWorkers; q++)
                            if (!B.fisher_busy[q]) {
                                B.fisher[q].release();
                            }
                    }
                    if (used_bytes(B) - mine + lane_bytes_after(*free_lane, evs, n) > B.budget) {
                        for (auto &r : batch)
                            finish(B, r, gwh_set_error(GWH_ERR_NOMEM, "request exceeds the HBM budget"));
                        batch.clear();
                    }
                }
            }
            if (!batch.empty()) {
                free_lane->batch = batch;
                int rc = launch_batch(B, *free_lane);
                if (rc != GWH_OK) {
                    (void)hipStreamSynchronize(free_lane->stream);
                    for (auto &r : free_lane->batch) finish(B, r, rc);
                    free_lane->batch.clear();
                } else {
                    free_lane->busy = true;
                }
            }
        }
        if (batch.empty()) {
            /* admission wait: block on the busy lanes' completion (or let a Fisher batch finish) */
            bool any = false;
            for (Lane &L : B.lanes)
                if (L.busy) {
                    (void)hipEventSynchronize(L.done);
                    reap(B, L);
                    any = true;
                }
            if (!any) std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    }
}

Broker *get(int id)
{
    std::lock_guard<std::mutex> lk(g_brokers_mu);
    auto it = g_brokers.find(id);
    return it == g_brokers.end() ? nullptr : it->second;
}

int submit(Broker *B, const std::shared_ptr<Request> &r, long long *ticket)
{
    std::lock_guard<std::mutex> lk(B->m);
    if (B->stop) return gwh_set_error(GWH_ERR_INVALID, "broker is shutting down");
    r->ticket = B->next_ticket++;
    r->submitted = clk::now();
    B->all[r->ticket] = r;
    *ticket = r->ticket;
    if (r->kind == REQ_FISHER) {
        B->fisher_queue.push_back(r);
        B->cv_fisher.notify_one();
    } else {
        B->queue.push_back(r);
        B->cv_work.notify_one();
    }
    return GWH_OK;
}

} // namespace

extern "C" {

int gwh_broker_create(size_t hbm_budget_bytes, int max_batch_instances, int *broker)
{
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (!broker || max_batch_instances < 1 || max_batch_instances > 65535)
        return gwh_set_error(GWH_ERR_INVALID, "max_batch_instances must be 1..65535");
    Broker *B = new Broker;
    B->budget = hbm_budget_bytes;
    B->max_n = max_batch_instances;
    for (Lane &L : B->lanes) {
        if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) {
            delete B;
            return gwh_set_error(GWH_ERR_HIP, "stream creation failed");
        }
    }
    for (int w = 0; w < kFisherWorkers; w++)
        if (hipStreamCreateWithFlags(&B->fisher[w].stream, hipStreamNonBlocking) != hipSuccess) {
            delete B;
            return gwh_set_error(GWH_ERR_HIP, "stream creation failed");
        }
    B->th = std::thread(dispatcher, B);
    for (int w = 0; w < kFisherWorkers; w++) B->fisher_th[w] = std::thread(fisher_worker, B, w);
    std::lock_guard<std::mutex> lk(g_brokers_mu);
    *broker = g_next_broker++;
    g_brokers[*broker] = B;
    return GWH_OK;
}

int gwh_broker_destroy(int broker)
{
    Broker *B;
    {
        std::lock_guard<std::mutex> lk(g_brokers_mu);
        auto it = g_brokers.find(broker);
        if (it == g_brokers.end()) return gwh_set_error(GWH_ERR_INVALID, "unknown broker");
        B = it->second;
        g_brokers.erase(it);
    }
    {
        std::lock_guard<std::mutex> lk(B->m);
        B->stop = true;
        B->cv_work.notify_all();
        B->cv_fisher.notify_all();
    }
    B->th.join(); /* drains the queue and the in-flight batches first */
    for (int w = 0; w < kFisherWorkers; w++) {
        B->fisher_th[w].join();
        FisherScratch &F = B->fisher[w];
        F.release(true);
        (void)hipStreamDestroy(F.stream);
    }
    /* tickets never collected release their events */
    for (auto &kv : B->all) unpin_events(*kv.second);
    for (Lane &L : B->lanes) {
        (void)hipStreamSynchronize(L.stream);
        (void)hipEventDestroy(L.done);
        (void)hipStreamDestroy(L.stream);
        if (L.h_params) (void)hipHostFree(L.h_params);
        if (L.h_events) (void)hipHostFree(L.h_events);
        if (L.h_logl) (void)hipHostFree(L.h_logl);
        for (DevBuf *d : {&L.scratch.coef, &L.scratch.partial, &L.scratch.active, &L.scratch.work, &L.scratch.wcount,
                          &L.d_params, &L.d_events, &L.d_logl})
            if (d->p) (void)hipFree(d->p);
    }
    delete B;
    return GWH_OK;
}

int gwh_broker_submit_loglike(int broker, const int *events, int event0, int model, int precision,
                              const gwh_params *params, int n, double *logl, long long *ticket)
{
    Broker *B = get(broker);
    if (!B) return gwh_set_error(GWH_ERR_INVALID, "unknown broker");
    if (!params || !logl || !ticket || n < 1) return gwh_set_error(GWH_ERR_INVALID, "bad arguments");
    if (!gwh_model_valid(model)) return gwh_set_error(GWH_ERR_UNSUPPORTED, "bad model");
    if (precision != GWH_FP64 && precision != GWH_MIXED) return gwh_set_error(GWH_ERR_INVALID, "bad precision");
    if (n > B->max_n) return gwh_set_error(GWH_ERR_INVALID, "request larger than the broker's batch limit");
    Engine &E = gwh_engine();
    auto r = std::make_shared<Request>();
    r->kind = REQ_LOGLIKE;
    r->model = model;
    r->precision = precision;
    r->n = n;
    r->params.assign(params, params + n);
    r->events.resize(n);
    for (int i = 0; i < n; i++) r->events[i] = events ? events[i] : event0;
    {
        std::lock_guard<std::recursive_mutex> lk(E.mu);
        int rc = pin_events(E, *r, r->events);
        if (rc != GWH_OK) return rc;
    }
    r->out = logl;
    int rc = submit(B, r, ticket);
    if (rc != GWH_OK) unpin_events(*r);
    return rc;
}

int gwh_broker_submit_fisher(int broker, int event, int model, int basis, int precision, const gwh_params *params,
                             int n, int dim, double epsilon, double *out, double *sigma, long long *ticket)
{
    Broker *B = get(broker);
    if (!B) return gwh_set_error(GWH_ERR_INVALID, "unknown broker");
    if (!params || !out || !ticket || n < 1) return gwh_set_error(GWH_ERR_INVALID, "bad arguments");
    auto r = std::make_shared<Request>();
    r->kind = REQ_FISHER;
    r->model = model;
    r->precision = precision;
    r->basis = basis;
    r->dim = dim;
    r->eps = epsilon;
    r->n = n;
    r->params.assign(params, params + n);
    r->events.assign(1, event);
    r->out = out;
    r->sigma = sigma;
    {
        Engine &E = gwh_engine();
        std::lock_guard<std::recursive_mutex> lk(E.mu);
        int rc = pin_events(E, *r, r->events);
        if (rc != GWH_OK) return rc;
    }
    int rc = submit(B, r, ticket);
    if (rc != GWH_OK) unpin_events(*r);
    return rc;
}

int gwh_broker_wait(int broker, long long ticket, double timeout_ms, double *latency_us)
{
    Broker *B = get(broker);
    if (!B) return gwh_set_error(GWH_ERR_INVALID, "unknown broker");
    std::unique_lock<std::mutex> lk(B->m);
    auto it = B->all.find(ticket);
    if (it == B->all.end()) return gwh_set_error(GWH_ERR_INVALID, "unknown ticket");
    auto r = it->second;
    auto ready = [&] { return r->status != 0; };
    if (timeout_ms < 0)
        r->done.wait(lk, ready);
    else if (!r->done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
        return gwh_set_error(GWH_ERR_BUSY, "ticket still pending");
    if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r->finished - r->submitted).count();
    B->all.erase(ticket);
    lk.unlock();
    unpin_events(*r); /* the request no longer reads its events */
    return r->status;
}

/* stats[0..8]: batches, instances, Fisher requests, most requests merged into one batch, admission waits,
 * peak device bytes (resident events + broker scratch), requests pending, Fisher batches, most Fisher requests
 * merged into one Fisher batch */
int gwh_broker_stats(int broker, double *stats)
{
    Broker *B = get(broker);
    if (!B || !stats) return gwh_set_error(GWH_ERR_INVALID, "unknown broker");
    std::lock_guard<std::mutex> lk(B->m);
    stats[0] = (double)B->batches;
    stats[1] = (double)B->instances;
    stats[2] = (double)B->fishers;
    stats[3] = (double)B->max_merged;
    stats[4] = (double)B->admission_waits;
    stats[5] = (double)B->peak_bytes;
    stats[6] = (double)B->all.size();
    stats[7] = (double)B->fisher_batches;
    stats[8] = (double)B->max_fisher_merged;
    return GWH_OK;
}

} /* extern "C" */
