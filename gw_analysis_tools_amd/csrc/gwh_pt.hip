/*
 * gwh_pt.hip -- sampler steps as replayable HIP graphs.
 *
 *   likelihood step   the batched likelihood (setup, per-bin tiles, finalize) on fixed device buffers,
 *                     captured once; every replay reads the caller's parameter buffer and writes its
 *                     log-likelihood buffer (for samplers that make their own proposals)
 *   tempering step    a whole parallel-tempering step of n_temps x n_chains chains on the device:
 *                     Gaussian proposal, prior box, likelihood of every proposal, Metropolis-Hastings
 *                     decision, adjacent-temperature swaps along every ladder; the step counter lives in
 *                     device memory, so one captured graph serves every step with no host round trip
 *
 * Reference: src/mcmc_sampler_internals.cpp:35-145 (mcmc_step), :364-409 (gaussian_step),
 * :1086-1115 (chain_swap), :1121-1180 (single_chain_swap).
 *
 * Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), key = the seed's two 32-bit halves,
 * counter = (instance or ladder pair, step low word, step high word, stream tag).  Per instance and step:
 * block A = tag 0, block B = tag 1; the proposal dimension is (A0 * dim) >> 32, the normal deviate is
 * Box-Muller sqrt(-2 ln(1 - u53(A1, A2))) cos(2 pi u53(A3, B0)), log U = ln u53(B1, B2); the swap uniform of
 * pair (k, chain c) is u53(C0, C1) of block (k * n_chains + c, step, tag 2).  u53(a, b) = ((a >> 5) 2^26 +
 * (b >> 6)) 2^-53.
 *
 * Mixed precision with the guard band: a Metropolis-Hastings decision whose margin (ll' - ll) / T - log U lies
 * within 2 tol / T of zero, and a swap whose margin lies within 2 tol |1/T2 - 1/T1|, is re-decided with FP64
 * likelihoods evaluated inside the same graph (a count-limited FP64 launch over the compacted states), so the
 * decision sequence equals the all-FP64 run's whenever the mixed likelihood error stays within tol.
 * A ladder with a swap inside the band is re-run with FP64 values for all of its states (the ladder's later
 * decisions depend on the earlier ones).
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gwh_basis.h"
#include "gwh_engine.h"

#define GWH_PT_MAXDIM 15
#define GWH_PT_MAXTEMPS 64
#define GWH_PT_MAXCHAINS 1024

/* ------------------------------------------------------------------------- */
/* Philox4x32-10                                                              */
/* ------------------------------------------------------------------------- */

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 o;
        o.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        o.y = (uint32_t)p1;
        o.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        o.w = (uint32_t)p0;
        c = o;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------- */
/* device state of one tempering run                                          */
/* ------------------------------------------------------------------------- */

struct PtArgs {
    int n, nt, nc, dim, basis, swap_every, guard;
    double tol;
    uint32_t key0, key1;
    const double *temps;   /* nt */
    const double *width;   /* n x dim: Gaussian step width of every chain and dimension */
    const double *lower, *upper; /* dim: prior box */
    const gwh_params *tmpl; /* non-sampled fields */
    double *theta, *ll;    /* current states */
    double *prop, *llp;    /* this step's proposals and their log-likelihoods */
    double *logu;          /* n: this step's log U */
    double *tmp;           /* n x dim: ladder permutation staging */
    gwh_params *params;    /* n: records of the proposals */
    int *mask;             /* n: proposal inside the prior box */
    int *acc;              /* n: 1 accepted, 0 rejected, 2 pending FP64 re-decision */
    int *swp;              /* (nt-1) x nc: pair swapped this step */
    long long *acc_count, *swap_count; /* n, (nt-1) x nc */
    unsigned long long *guard_evals;
    long long *step;
    /* guard band */
    int *gcount;           /* [0]: MH states, [1]: ladder states */
    int *gslot;            /* n: slot of instance i's pair in gparams */
    int *lslot;            /* nc: slot of ladder c's states in lparams, -1 none */
    gwh_params *gparams;   /* 2n */
    gwh_params *lparams;   /* n */
    double *gll, *lll;     /* FP64 log-likelihoods of gparams, lparams */
};

__device__ __forceinline__ gwh_params pt_record(const PtArgs &a, const double *v)
{
    gwh_params p = *a.tmpl;
    repack(v, a.basis, p);
    return p;
}

__device__ __forceinline__ double pt_T(const PtArgs &a, int i) { return a.temps[i / a.nc]; }

/* records of the initial states (for their likelihoods) */
__global__ void gwh_pt_init_kernel(PtArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    a.params[i] = pt_record(a, a.theta + (size_t)i * a.dim);
    a.mask[i] = 1;
}

/* gaussian_step (mcmc_sampler_internals.cpp:364-409): one dimension chosen uniformly, moved by
 * width * N(0, 1); the prior box decides whether the likelihood is needed at all (mcmc_step :83-92) */
__global__ __launch_bounds__(256) void gwh_pt_propose_kernel(PtArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        a.gcount[0] = 0;
        a.gcount[1] = 0;
    }
    if (i >= a.n) return;
    const long long s = *a.step;
    const uint32_t slo = (uint32_t)s, shi = (uint32_t)((unsigned long long)s >> 32);
    const U4 A = philox4x32_10(U4{(uint32_t)i, slo, shi, 0u}, a.key0, a.key1);
    const U4 B = philox4x32_10(U4{(uint32_t)i, slo, shi, 1u}, a.key0, a.key1);
    const int beta = (int)(((uint64_t)A.x * (uint64_t)a.dim) >> 32);
    const double u1 = u53(A.y, A.z), u2 = u53(A.w, B.x);
    const double z = sqrt(-2.0 * log(1.0 - u1)) * cospi(2.0 * u2);
    a.logu[i] = log(u53(B.y, B.z));
    const double *th = a.theta + (size_t)i * a.dim;
    double *pr = a.prop + (size_t)i * a.dim;
    /* fixed trip count: every index of v is a compile-time constant, so v lives in registers */
    double v[GWH_PT_MAXDIM], t_beta = 0;
    int ok = 1;
#pragma unroll
    for (int d = 0; d < GWH_PT_MAXDIM; d++) {
        v[d] = 0;
        if (d < a.dim) {
            double x = th[d];
            if (d == beta) {
                t_beta = x;
                x = __dadd_rn(__dmul_rn(a.width[(size_t)i * a.dim + d], z), x);
            }
            v[d] = x;
            pr[d] = x;
            ok &= (x >= a.lower[d]) & (x <= a.upper[d]);
        }
    }
    a.mask[i] = ok;
    if (!ok) { /* outside the prior box: the record of the current state (its likelihood is not used) */
#pragma unroll
        for (int d = 0; d < GWH_PT_MAXDIM; d++)
            if (d == beta) v[d] = t_beta;
    }
    /* a.params[i] holds the template's other fields since the init kernel: only the sampled ones change (the
     * sampling bases' repack reads no field of the record) */
    repack(v, a.basis, a.params[i]);
}

/* Metropolis-Hastings (mcmc_sampler_internals.cpp:96-142): reject iff (ll' - ll) / T < log U; a proposal
 * outside the prior box or with a NaN likelihood is rejected */
__device__ __forceinline__ void pt_apply(const PtArgs &a, int i, bool accept)
{
    a.acc[i] = accept;
    if (!accept) return;
    for (int d = 0; d < a.dim; d++) a.theta[(size_t)i * a.dim + d] = a.prop[(size_t)i * a.dim + d];
    a.ll[i] = a.llp[i];
    a.acc_count[i]++;
}

__device__ __forceinline__ bool mh_accept(double ratio, double logu)
{
    return ratio > -INFINITY && !(ratio < logu);
}

__global__ void gwh_pt_accept_kernel(PtArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double T = pt_T(a, i), lu = a.logu[i];
    double lp = a.llp[i];
    if (!a.mask[i]) {
        lp = __longlong_as_double(0x7ff8000000000000ll);
        a.llp[i] = lp;
    }
    const double ratio = (lp == lp) ? (-a.ll[i] + lp) / T : -INFINITY;
    const double margin = ratio - lu;
    if (a.guard && ratio > -INFINITY && isfinite(margin) && fabs(margin) < 2.0 * a.tol / T) {
        const int slot = atomicAdd(&a.gcount[0], 2);
        a.gslot[i] = slot;
        a.gparams[slot] = a.params[i];
        a.gparams[slot + 1] = pt_record(a, a.theta + (size_t)i * a.dim);
        a.acc[i] = 2;
        return;
    }
    pt_apply(a, i, mh_accept(ratio, lu));
}

__global__ void gwh_pt_accept_guard_kernel(PtArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n || a.acc[i] != 2) return;
    const int slot = a.gslot[i];
    const double T = pt_T(a, i);
    const double l1 = a.gll[slot], l0 = a.gll[slot + 1];
    const double ratio = (l1 == l1 && l0 == l0) ? (-l0 + l1) / T : -INFINITY;
    atomicAdd(a.guard_evals, 2ull);
    pt_apply(a, i, mh_accept(ratio, a.logu[i]));
}

/* chain_swap / single_chain_swap (mcmc_sampler_internals.cpp:1086-1180): pairs (k, k+1) of one ladder in
 * increasing k; reject iff exp((ll1 - ll2) / T2 - (ll1 - ll2) / T1) < U.  perm[k] = the temperature index
 * whose state (at the start of the swaps) sits at temperature k. */
__device__ int pt_ladder(const PtArgs &a, int c, long long s, const double *ll64, int *perm, int *flag)
{
    const uint32_t slo = (uint32_t)s, shi = (uint32_t)((unsigned long long)s >> 32);
    for (int k = 0; k < a.nt; k++) perm[k] = k;
    for (int k = 0; k + 1 < a.nt; k++) {
        const double T1 = a.temps[k], T2 = a.temps[k + 1];
        flag[k] = 0;
        if (T1 == T2) continue;
        const U4 C = philox4x32_10(U4{(uint32_t)(k * a.nc + c), slo, shi, 2u}, a.key0, a.key1);
        const double alpha = u53(C.x, C.y);
        const double l1 = a.ll[perm[k] * a.nc + c], l2 = a.ll[perm[k + 1] * a.nc + c];
        double pw = (l1 - l2) / T2 - (l1 - l2) / T1;
        if (a.guard && alpha > 0 && fabs(pw - log(alpha)) < 2.0 * a.tol * fabs(1.0 / T2 - 1.0 / T1)) {
            if (!ll64) return -1; /* FP64 values of this ladder's states are needed */
            const double m1 = ll64[perm[k]], m2 = ll64[perm[k + 1]];
            pw = (m1 - m2) / T2 - (m1 - m2) / T1;
        }
        if (!(exp(pw) < alpha)) {
            const int t = perm[k];
            perm[k] = perm[k + 1];
            perm[k + 1] = t;
            flag[k] = 1;
        }
    }
    return 0;
}

__device__ void pt_permute(const PtArgs &a, int c, const int *perm, const int *flag)
{
    double lv[GWH_PT_MAXTEMPS];
    for (int k = 0; k < a.nt; k++) {
        const int i = k * a.nc + c;
        lv[k] = a.ll[i];
        for (int d = 0; d < a.dim; d++) a.tmp[(size_t)i * a.dim + d] = a.theta[(size_t)i * a.dim + d];
    }
    for (int k = 0; k < a.nt; k++) {
        const int i = k * a.nc + c, j = perm[k] * a.nc + c;
        a.ll[i] = lv[perm[k]];
        for (int d = 0; d < a.dim; d++) a.theta[(size_t)i * a.dim + d] = a.tmp[(size_t)j * a.dim + d];
    }
    for (int k = 0; k + 1 < a.nt; k++) {
        a.swp[k * a.nc + c] = flag[k];
        a.swap_count[k * a.nc + c] += flag[k];
    }
}

/* one block of n_chains threads; the last kernel of a step advances the step counter */
__global__ __launch_bounds__(1024) void gwh_pt_swap_kernel(PtArgs a, int last)
{
    const int c = threadIdx.x;
    const long long s = *a.step;
    const bool swap_step = a.swap_every > 0 && (s + 1) % a.swap_every == 0;
    if (c < a.nc) {
        a.lslot[c] = -1;
        int perm[GWH_PT_MAXTEMPS], flag[GWH_PT_MAXTEMPS];
        for (int k = 0; k + 1 < a.nt; k++) a.swp[k * a.nc + c] = 0;
        if (swap_step) {
            if (pt_ladder(a, c, s, nullptr, perm, flag) == 0) {
                pt_permute(a, c, perm, flag);
            } else {
                const int slot = atomicAdd(&a.gcount[1], a.nt);
                a.lslot[c] = slot;
                for (int k = 0; k < a.nt; k++)
                    a.lparams[slot + k] = pt_record(a, a.theta + (size_t)(k * a.nc + c) * a.dim);
            }
        }
    }
    __syncthreads();
    if (last && threadIdx.x == 0) *a.step = s + 1;
}

__global__ __launch_bounds__(1024) void gwh_pt_swap_guard_kernel(PtArgs a)
{
    const int c = threadIdx.x;
    const long long s = *a.step;
    if (c < a.nc && a.lslot[c] >= 0) {
        int perm[GWH_PT_MAXTEMPS], flag[GWH_PT_MAXTEMPS];
        pt_ladder(a, c, s, a.lll + a.lslot[c], perm, flag);
        pt_permute(a, c, perm, flag);
        atomicAdd(a.guard_evals, (unsigned long long)a.nt);
    }
    __syncthreads();
    if (threadIdx.x == 0) *a.step = s + 1;
}

/* ------------------------------------------------------------------------- */
/* host side                                                                  */
/* ------------------------------------------------------------------------- */

namespace {

struct StepGraph {
    std::mutex mu;
    bool dead = false;
    int kind = 0; /* 0 likelihood step, 1 tempering run */
    std::vector<int> pinned;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Scratch scr;
    /* tempering run */
    int steps_per_graph = 1;
    PtArgs a;
    DevBuf state;
    int event = -1, model = 0, precision = 0;
};

std::mutex g_mu;
std::map<int, std::shared_ptr<StepGraph>> g_graphs;
int g_next = 1;

std::shared_ptr<StepGraph> find_graph(int id, int kind)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_graphs.find(id);
    if (it == g_graphs.end() || it->second->kind != kind) return nullptr;
    return it->second;
}

int register_graph(std::shared_ptr<StepGraph> g)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int id = g_next++;
    g_graphs[id] = g;
    return id;
}

/* release device resources (engine lock held by the caller) */
void free_graph(StepGraph &g)
{
    Engine &E = gwh_engine();
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.stream) (void)hipStreamDestroy(g.stream);
    g.exec = nullptr;
    g.graph = nullptr;
    g.stream = nullptr;
    g.scr.release();
    g.state.release();
    for (int e : g.pinned)
        if (e >= 0 && e < GWH_MAX_EVENTS && E.events[e].pins > 0) E.events[e].pins--;
    g.pinned.clear();
}

int pin_events(StepGraph &g, const int *events, int n, int event0)
{
    Engine &E = gwh_engine();
    std::vector<char> seen(GWH_MAX_EVENTS, 0);
    auto add = [&](int e) -> int {
        if (e < 0 || e >= GWH_MAX_EVENTS || !E.events[e].live)
            return gwh_set_error(GWH_ERR_INVALID, "unknown event " + std::to_string(e));
        if (!seen[e]) {
            seen[e] = 1;
            E.events[e].pins++;
            g.pinned.push_back(e);
        }
        return GWH_OK;
    };
    if (events) {
        for (int i = 0; i < n; i++) {
            int rc = add(events[i]);
            if (rc != GWH_OK) return rc;
        }
        return GWH_OK;
    }
    return add(event0);
}

/* capture what `body` enqueues on g.stream into g.graph / g.exec */
template <class F> int capture(StepGraph &g, F body)
{
    GWH_HIP(hipStreamBeginCapture(g.stream, hipStreamCaptureModeRelaxed));
    int rc = body();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g.stream, &graph);
    if (rc != GWH_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return gwh_set_error(GWH_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    g.graph = graph;
    GWH_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    return GWH_OK;
}

inline dim3 blocks_of(int n) { return dim3((n + 255) / 256); }

} // namespace

/* ---- likelihood step ---- */

int gwh_loglike_graph_create(const int *events, const int *d_events, int events0, int model, int precision,
                             const void *d_params, int n, double *d_logl, int *graph)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (!graph || !d_params || !d_logl || n <= 0) return gwh_set_error(GWH_ERR_INVALID, "bad argument");
    if ((events == nullptr) != (d_events == nullptr))
        return gwh_set_error(GWH_ERR_INVALID, "events and d_events are given together or not at all");
    auto g = std::make_shared<StepGraph>();
    g->kind = 0;
    int rc = pin_events(*g, events, n, events0);
    if (rc == GWH_OK && hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        rc = gwh_set_error(GWH_ERR_HIP, "hipStreamCreateWithFlags failed");
    LaunchExtra x;
    x.reserve_only = true;
    if (rc == GWH_OK)
        rc = gwh_loglike_launch(d_events, events ? events[0] : events0, model, precision, (const gwh_params *)d_params,
                                n, d_logl, g->stream, false, nullptr, events, 0, &g->scr, &x);
    if (rc == GWH_OK) {
        x.reserve_only = false;
        x.capturing = true;
        rc = capture(*g, [&]() {
            return gwh_loglike_launch(d_events, events ? events[0] : events0, model, precision,
                                      (const gwh_params *)d_params, n, d_logl, g->stream, false, nullptr, events, 0,
                                      &g->scr, &x);
        });
    }
    if (rc != GWH_OK) {
        free_graph(*g);
        return rc;
    }
    *graph = register_graph(g);
    return GWH_OK;
}

int gwh_graph_launch(int graph, void *stream)
{
    auto g = find_graph(graph, 0);
    if (!g) return gwh_set_error(GWH_ERR_INVALID, "unknown graph");
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->dead) return gwh_set_error(GWH_ERR_INVALID, "unknown graph");
    GWH_HIP(hipGraphLaunch(g->exec, stream ? (hipStream_t)stream : g->stream));
    return GWH_OK;
}

int gwh_graph_synchronize(int graph)
{
    auto g = find_graph(graph, 0);
    if (!g) g = find_graph(graph, 1);
    if (!g) return gwh_set_error(GWH_ERR_INVALID, "unknown graph");
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->dead) return gwh_set_error(GWH_ERR_INVALID, "unknown graph");
    GWH_HIP(hipStreamSynchronize(g->stream));
    return GWH_OK;
}

static int destroy_graph(int id, int kind)
{
    std::shared_ptr<StepGraph> g;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_graphs.find(id);
        if (it == g_graphs.end() || it->second->kind != kind) return gwh_set_error(GWH_ERR_INVALID, "unknown graph");
        g = it->second;
        g_graphs.erase(it);
    }
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    std::lock_guard<std::mutex> lk2(g->mu);
    g->dead = true;
    free_graph(*g);
    return GWH_OK;
}

int gwh_graph_destroy(int graph) { return destroy_graph(graph, 0); }

/* ---- tempering run ---- */

int gwh_pt_create(int event, const gwh_pt_config *cfg, const double *temps, const double *widths,
                  const double *lower, const double *upper, const gwh_params *tmpl, const double *theta0, int *pt)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (!cfg || !temps || !widths || !lower || !upper || !tmpl || !theta0 || !pt)
        return gwh_set_error(GWH_ERR_INVALID, "null argument");
    const int nt = cfg->n_temps, nc = cfg->n_chains;
    if (nt < 1 || nt > GWH_PT_MAXTEMPS) return gwh_set_error(GWH_ERR_INVALID, "n_temps must be 1..64");
    if (nc < 1 || nc > GWH_PT_MAXCHAINS) return gwh_set_error(GWH_ERR_INVALID, "n_chains must be 1..1024");
    const int n = nt * nc;
    if (2 * n > 65535) return gwh_set_error(GWH_ERR_INVALID, "at most 32767 chains in all");
    if (cfg->basis != GWH_BASIS_PHENOMD_MCMC && cfg->basis != GWH_BASIS_PV2_MCMC)
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "sampling basis must be GWH_BASIS_PHENOMD_MCMC or GWH_BASIS_PV2_MCMC");
    if (cfg->model != GWH_IMRPHENOMD && cfg->model != GWH_IMRPHENOMPV2)
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "model must be GWH_IMRPHENOMD or GWH_IMRPHENOMPV2");
    if (cfg->precision != GWH_FP64 && cfg->precision != GWH_MIXED)
        return gwh_set_error(GWH_ERR_INVALID, "bad precision");
    if (cfg->swap_every < 0 || cfg->steps_per_graph < 1 || cfg->steps_per_graph > 1000)
        return gwh_set_error(GWH_ERR_INVALID, "swap_every >= 0 and 1 <= steps_per_graph <= 1000");
    for (int k = 0; k < nt; k++)
        if (!(temps[k] > 0) || !std::isfinite(temps[k]))
            return gwh_set_error(GWH_ERR_INVALID, "temperatures must be positive and finite");
    const int dim = basis_dim(cfg->basis);
    const bool guard = cfg->guard && cfg->precision == GWH_MIXED;
    if (guard && !(cfg->tol > 0)) return gwh_set_error(GWH_ERR_INVALID, "guard band needs tol > 0");

    auto g = std::make_shared<StepGraph>();
    g->kind = 1;
    g->steps_per_graph = cfg->steps_per_graph;
    g->event = event;
    g->model = cfg->model;
    g->precision = cfg->precision;
    int rc = pin_events(*g, nullptr, 0, event);
    if (rc != GWH_OK) return rc;
    auto fail = [&](int code) {
        free_graph(*g);
        return code;
    };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(gwh_set_error(GWH_ERR_HIP, "hipStreamCreateWithFlags failed"));

    /* one allocation: doubles, then records, then 8-byte counters, then ints */
    const size_t nd = (size_t)nt + (size_t)n * dim + 2 * dim + 1 /* tmpl slot (aligned) */ + 2 * (size_t)n * dim +
                      3 * (size_t)n + (size_t)n * dim + 2 * (size_t)n + (size_t)n;
    const size_t nrec = (size_t)n + 2 * (size_t)n + (size_t)n + 1;
    const size_t ncnt = (size_t)n + (size_t)(nt - 1) * nc + 2;
    const size_t nint = (size_t)n * 3 + (size_t)(nt - 1) * nc + 2 + (size_t)n + (size_t)nc;
    const size_t bytes = nd * 8 + nrec * sizeof(gwh_params) + ncnt * 8 + nint * 4 + 256;
    if ((rc = g->state.ensure(bytes)) != GWH_OK) return fail(rc);
    char *base = (char *)g->state.p;
    size_t off = 0;
    auto take = [&](size_t b) {
        char *p = base + off;
        off += (b + 15) & ~(size_t)15;
        return p;
    };
    PtArgs &a = g->a;
    a.n = n;
    a.nt = nt;
    a.nc = nc;
    a.dim = dim;
    a.basis = cfg->basis;
    a.swap_every = cfg->swap_every;
    a.guard = guard;
    a.tol = cfg->tol;
    a.key0 = (uint32_t)cfg->seed;
    a.key1 = (uint32_t)(cfg->seed >> 32);
    double *d_temps = (double *)take(nt * 8);
    double *d_width = (double *)take((size_t)n * dim * 8);
    double *d_lower = (double *)take(dim * 8), *d_upper = (double *)take(dim * 8);
    a.temps = d_temps;
    a.width = d_width;
    a.lower = d_lower;
    a.upper = d_upper;
    a.theta = (double *)take((size_t)n * dim * 8);
    a.prop = (double *)take((size_t)n * dim * 8);
    a.ll = (double *)take(n * 8);
    a.llp = (double *)take(n * 8);
    a.logu = (double *)take(n * 8);
    a.tmp = (double *)take((size_t)n * dim * 8);
    a.gll = (double *)take(2 * (size_t)n * 8);
    a.lll = (double *)take((size_t)n * 8);
    gwh_params *d_tmpl = (gwh_params *)take(sizeof(gwh_params));
    a.tmpl = d_tmpl;
    a.params = (gwh_params *)take((size_t)n * sizeof(gwh_params));
    a.gparams = (gwh_params *)take(2 * (size_t)n * sizeof(gwh_params));
    a.lparams = (gwh_params *)take((size_t)n * sizeof(gwh_params));
    a.acc_count = (long long *)take((size_t)n * 8);
    a.swap_count = (long long *)take((size_t)std::max(nt - 1, 1) * nc * 8);
    a.guard_evals = (unsigned long long *)take(8);
    a.step = (long long *)take(8);
    a.mask = (int *)take((size_t)n * 4);
    a.acc = (int *)take((size_t)n * 4);
    a.gslot = (int *)take((size_t)n * 4);
    a.swp = (int *)take((size_t)std::max(nt - 1, 1) * nc * 4);
    a.gcount = (int *)take(2 * 4);
    a.lslot = (int *)take((size_t)nc * 4);
    if (off > g->state.cap) return fail(gwh_set_error(GWH_ERR_INVALID, "internal: state layout overflow"));

    hipStream_t s = g->stream;
    auto up = [&](void *dst, const void *src, size_t b) {
        return hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, s) == hipSuccess;
    };
    bool ok = up(d_temps, temps, nt * 8) && up(d_width, widths, (size_t)n * dim * 8) &&
              up(d_lower, lower, dim * 8) && up(d_upper, upper, dim * 8) &&
              up(a.theta, theta0, (size_t)n * dim * 8) && up(d_tmpl, tmpl, sizeof(gwh_params));
    ok = ok && hipMemsetAsync(a.acc_count, 0, (size_t)n * 8, s) == hipSuccess &&
         hipMemsetAsync(a.swap_count, 0, (size_t)std::max(nt - 1, 1) * nc * 8, s) == hipSuccess &&
         hipMemsetAsync(a.guard_evals, 0, 8, s) == hipSuccess && hipMemsetAsync(a.step, 0, 8, s) == hipSuccess &&
         hipMemsetAsync(a.acc, 0, (size_t)n * 4, s) == hipSuccess &&
         hipMemsetAsync(a.swp, 0, (size_t)std::max(nt - 1, 1) * nc * 4, s) == hipSuccess &&
         hipMemsetAsync(a.lslot, 0xff, (size_t)nc * 4, s) == hipSuccess &&
         hipMemsetAsync(a.prop, 0, (size_t)n * dim * 8, s) == hipSuccess &&
         hipMemsetAsync(a.llp, 0, (size_t)n * 8, s) == hipSuccess;
    if (!ok) return fail(gwh_set_error(GWH_ERR_HIP, "initial copies failed"));

    /* scratch sized for the largest launch before anything is captured */
    LaunchExtra x;
    x.reserve_only = true;
    const int P = cfg->precision, M = cfg->model;
    if ((rc = gwh_loglike_launch(nullptr, event, M, P, a.params, n, a.llp, s, false, nullptr, nullptr, 0, &g->scr,
                                 &x)) != GWH_OK)
        return fail(rc);
    if (guard && (rc = gwh_loglike_launch(nullptr, event, M, GWH_FP64, a.gparams, 2 * n, a.gll, s, false, nullptr,
                                          nullptr, 0, &g->scr, &x)) != GWH_OK)
        return fail(rc);

    /* likelihoods of the initial states (eager) */
    x.reserve_only = false;
    hipLaunchKernelGGL(gwh_pt_init_kernel, blocks_of(n), dim3(256), 0, s, a);
    if ((rc = gwh_loglike_launch(nullptr, event, M, P, a.params, n, a.ll, s, false, nullptr, nullptr, 0, &g->scr,
                                 &x)) != GWH_OK)
        return fail(rc);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(gwh_set_error(GWH_ERR_HIP, "initial likelihoods failed"));

    x.capturing = true;
    const int nsw = std::max(nc, 64);
    rc = capture(*g, [&]() -> int {
        int r;
        for (int k = 0; k < g->steps_per_graph; k++) {
            hipLaunchKernelGGL(gwh_pt_propose_kernel, blocks_of(n), dim3(256), 0, s, a);
            LaunchExtra m = x;
            m.d_mask = a.mask;
            if ((r = gwh_loglike_launch(nullptr, event, M, P, a.params, n, a.llp, s, false, nullptr, nullptr, 0,
                                        &g->scr, &m)) != GWH_OK)
                return r;
            hipLaunchKernelGGL(gwh_pt_accept_kernel, blocks_of(n), dim3(256), 0, s, a);
            if (guard) {
                LaunchExtra c0 = x;
                c0.d_count = a.gcount;
                if ((r = gwh_loglike_launch(nullptr, event, M, GWH_FP64, a.gparams, 2 * n, a.gll, s, false, nullptr,
                                            nullptr, 0, &g->scr, &c0)) != GWH_OK)
                    return r;
                hipLaunchKernelGGL(gwh_pt_accept_guard_kernel, blocks_of(n), dim3(256), 0, s, a);
            }
            hipLaunchKernelGGL(gwh_pt_swap_kernel, dim3(1), dim3(nsw), 0, s, a, guard ? 0 : 1);
            if (guard) {
                LaunchExtra c1 = x;
                c1.d_count = a.gcount + 1;
                if ((r = gwh_loglike_launch(nullptr, event, M, GWH_FP64, a.lparams, n, a.lll, s, false, nullptr,
                                            nullptr, 0, &g->scr, &c1)) != GWH_OK)
                    return r;
                hipLaunchKernelGGL(gwh_pt_swap_guard_kernel, dim3(1), dim3(nsw), 0, s, a);
            }
            GWH_HIP(hipGetLastError());
        }
        return GWH_OK;
    });
    if (rc != GWH_OK) return fail(rc);
    *pt = register_graph(g);
    return GWH_OK;
}

int gwh_pt_run(int pt, int steps)
{
    auto g = find_graph(pt, 1);
    if (!g) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->dead) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    if (steps < 0 || steps % g->steps_per_graph != 0)
        return gwh_set_error(GWH_ERR_INVALID, "steps must be a non-negative multiple of steps_per_graph");
    for (int k = 0; k < steps / g->steps_per_graph; k++) GWH_HIP(hipGraphLaunch(g->exec, g->stream));
    return GWH_OK;
}

int gwh_pt_synchronize(int pt)
{
    auto g = find_graph(pt, 1);
    if (!g) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->dead) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    GWH_HIP(hipStreamSynchronize(g->stream));
    return GWH_OK;
}

static int pt_read(int pt, const std::vector<std::pair<void *, std::pair<const void *, size_t>>> &items)
{
    auto g = find_graph(pt, 1);
    if (!g) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->dead) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    GWH_HIP(hipStreamSynchronize(g->stream));
    for (auto &it : items)
        if (it.first) GWH_HIP(hipMemcpy(it.first, it.second.first, it.second.second, hipMemcpyDeviceToHost));
    return GWH_OK;
}

int gwh_pt_state(int pt, double *theta, double *ll, long long *step)
{
    auto g = find_graph(pt, 1);
    if (!g) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    const PtArgs &a = g->a;
    return pt_read(pt, {{theta, {a.theta, (size_t)a.n * a.dim * 8}}, {ll, {a.ll, (size_t)a.n * 8}},
                        {step, {a.step, 8}}});
}

int gwh_pt_last_step(int pt, double *prop, double *llp, int *accept, int *swapped)
{
    auto g = find_graph(pt, 1);
    if (!g) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    const PtArgs &a = g->a;
    return pt_read(pt, {{prop, {a.prop, (size_t)a.n * a.dim * 8}}, {llp, {a.llp, (size_t)a.n * 8}},
                        {accept, {a.acc, (size_t)a.n * 4}},
                        {swapped, {a.swp, (size_t)std::max(a.nt - 1, 0) * a.nc * 4}}});
}

int gwh_pt_counters(int pt, long long *accepts, long long *swaps, long long *guard_evals)
{
    auto g = find_graph(pt, 1);
    if (!g) return gwh_set_error(GWH_ERR_INVALID, "unknown tempering run");
    const PtArgs &a = g->a;
    return pt_read(pt, {{accepts, {a.acc_count, (size_t)a.n * 8}},
                        {swaps, {a.swap_count, (size_t)std::max(a.nt - 1, 0) * a.nc * 8}},
                        {guard_evals, {a.guard_evals, 8}}});
}

int gwh_pt_destroy(int pt) { return destroy_graph(pt, 1); }
