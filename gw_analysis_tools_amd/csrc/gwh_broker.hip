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
 *  - Fisher requests run one at a time on a worker thread with its own stream and scratch, concurrently
 *    with the likelihood batches.
 *
 * An instance's result does not depend on what it is batched with (per-event tiling, fixed-order sums),
 * so a request's results are bit-identical alone and in the mix.  Events referenced by pending
 * requests must stay alive until those requests complete.
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
    int status = 0; /* 0 pending, else the GWH status */
};

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
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::map<long long, std::shared_ptr<Request>> all;
    long long next_ticket = 1;
    bool stop = false;
    Lane lanes[2];
    std::thread th;
    /* Fisher requests: their own queue, worker thread, stream and scratch */
    std::deque<std::shared_ptr<Request>> fisher_queue;
    std::condition_variable cv_fisher;
    FisherScratch fisher;
    std::thread fisher_th;
    /* statistics */
    long long batches = 0, instances = 0, fishers = 0, max_merged = 0, admission_waits = 0;
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

/* device bytes a batch of n instances over `events` needs in one lane's scratch */
size_t batch_bytes(const std::vector<int> &events, int n)
{
    Engine &E = gwh_engine();
    size_t maxtiles = 1, tiles = 0;
    for (int i = 0; i < n; i++) {
        const Event &ev = E.events[events[i]];
        size_t t = (ev.N + GWH_TPB * ev.desc.bpl - 1) / (GWH_TPB * ev.desc.bpl);
        tiles += t;
        if (t > maxtiles) maxtiles = t;
    }
    return (size_t)n * (sizeof(InstCoef) + sizeof(gwh_params) + 2 * sizeof(int) + sizeof(double) +
                        maxtiles * GWH_PARTIAL_STRIDE * sizeof(double)) +
           tiles * sizeof(int2) + 64;
}

void finish(Broker &B, const std::shared_ptr<Request> &r, int status)
{
    r->finished = clk::now();
    std::lock_guard<std::mutex> lk(B.m);
    r->status = status;
    B.cv_done.notify_all();
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
    size_t used = resident_bytes() + lane_bytes(B.lanes[0]) + lane_bytes(B.lanes[1]);
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

void fisher_worker(Broker *bp)
{
    Broker &B = *bp;
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(B.m);
            B.cv_fisher.wait(lk, [&] { return B.stop || !B.fisher_queue.empty(); });
            if (B.fisher_queue.empty()) return; /* stop requested and nothing left */
            r = B.fisher_queue.front();
            B.fisher_queue.pop_front();
        }
        int st = gwh_fisher_batch_with(r->events[0], r->model, r->basis, r->precision, r->params.data(), r->n, r->dim,
                                       r->eps, r->out, r->sigma, B.fisher);
        {
            std::lock_guard<std::mutex> lk(B.m);
            B.fishers++;
        }
        finish(B, r, st);
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
        size_t need;
        {
            std::lock_guard<std::recursive_mutex> lk(E.mu);
            need = batch_bytes(evs, n);
            size_t used = resident_bytes() + lane_bytes(B.lanes[0]) + lane_bytes(B.lanes[1]);
            size_t mine = lane_bytes(*free_lane);
            if (used - mine + need > B.budget) {
                bool other_busy = B.lanes[0].busy || B.lanes[1].busy;
                if (other_busy) {
                    /* wait for in-flight work, then retry with the batch back at the queue head */
                    std::lock_guard<std::mutex> lq(B.m);
                    for (auto it = batch.rbegin(); it != batch.rend(); ++it) B.queue.push_front(*it);
                    B.admission_waits++;
                    batch.clear();
                } else if (resident_bytes() + need > B.budget) {
                    for (auto &r : batch) finish(B, r, gwh_set_error(GWH_ERR_NOMEM, "request exceeds the HBM budget"));
                    batch.clear();
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
            /* admission wait: block on the busy lanes' completion */
            for (Lane &L : B.lanes)
                if (L.busy) {
                    (void)hipEventSynchronize(L.done);
                    reap(B, L);
                }
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
    if (hipStreamCreateWithFlags(&B->fisher.stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return gwh_set_error(GWH_ERR_HIP, "stream creation failed");
    }
    B->th = std::thread(dispatcher, B);
    B->fisher_th = std::thread(fisher_worker, B);
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
    B->fisher_th.join();
    (void)hipStreamSynchronize(B->fisher.stream);
    for (DevBuf *d : {&B->fisher.params, &B->fisher.coef, &B->fisher.misc, &B->fisher.deriv, &B->fisher.gram,
                      &B->fisher.out, &B->fisher.sig, &B->fisher.raw})
        if (d->p) (void)hipFree(d->p);
    (void)hipStreamDestroy(B->fisher.stream);
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
    {
        std::lock_guard<std::recursive_mutex> lk(E.mu);
        for (int i = 0; i < n; i++) {
            int e = events ? events[i] : event0;
            if (e < 0 || e >= GWH_MAX_EVENTS || !E.events[e].live)
                return gwh_set_error(GWH_ERR_INVALID, "unknown event in request");
            r->events[i] = e;
        }
    }
    r->out = logl;
    return submit(B, r, ticket);
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
    return submit(B, r, ticket);
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
        B->cv_done.wait(lk, ready);
    else if (!B->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
        return gwh_set_error(GWH_ERR_BUSY, "ticket still pending");
    if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r->finished - r->submitted).count();
    B->all.erase(ticket);
    return r->status;
}

/* stats[0..6]: batches, instances, Fisher requests, most requests merged into one batch, admission waits,
 * peak device bytes (resident events + broker scratch), requests pending */
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
    return GWH_OK;
}

} /* extern "C" */
