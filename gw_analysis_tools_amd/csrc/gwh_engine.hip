/*
 * gwh_engine.hip -- runtime, resident events and the batched network
 * log-likelihood of IMRPhenomD / IMRPhenomPv2 on MI355X (gfx950).
 *
 * Hot path (one MCMC step):
 *   1. gwh_setup_kernel      FP64 InstCoef record of every instance: four waves per
 *                            64 instances compute independent parts, joined in LDS
 *   2. loglike kernel        (instance, bin-tile) work items, one block each
 *                            (gwh_tiles.h): in tile-major order for one event
 *                            (or events whose sizes are known only on the
 *                            device), or from flattened work lists for
 *                            multi-event batches; a tile is lanes x BPL bins,
 *                            both per event (autotuned); tiles above 0.2/M of
 *                            an ascending grid do nothing; the instance record
 *                            uniform (scalar loads); per detector accumulate
 *                            |h|^2 w/S and h* d w/S in FP64; wave shuffle + LDS
 *                            tree per item -> partials
 *   3. gwh_finalize_kernel   one wave per instance sums its items in a fixed
 *                            order (deterministic, batch-composition free)
 *
 * Precision modes: GWH_FP64 evaluates everything in FP64, sin/cos and exp through
 * correctly rounded tables and short series (gwh_trig.h).  GWH_MIXED differs only
 * where FP32 is faster on gfx950: with packed FP32 disabled, FP32 and FP64 FMAs
 * issue at the same rate, so only FP32 hardware transcendentals can pay --
 * the IMRPhenomD merger-ringdown exponential (FP32 exp2 with an FP64 exponent
 * split) and the IMRPhenomPv2 reciprocal square root's seed.  Phases, products
 * and sums stay FP64.
 *
 * Reference: src/mcmc_gw.cpp:2374-2489 (MCMC_likelihood_extrinsic),
 * :801-868 (Log_Likelihood_internal), src/waveform_util.cpp:152-184
 * (create_coherent_GW_detection_reuse_WF), src/IMRPhenomD.cpp:403-504.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gwh_coalesce.h"
#include "gwh_engine.h"
#include "gwh_tables.h"
#include "gwh_tiles.h"

/* ------------------------------------------------------------------------- */
/* runtime                                                                    */
/* ------------------------------------------------------------------------- */

static thread_local std::string g_err;

int gwh_set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

Engine &gwh_engine()
{
    static Engine *e = new Engine(); /* lives until process exit */
    return *e;
}

int DevBuf::ensure(size_t bytes)
{
    if (bytes <= cap) return GWH_OK;
    if (fixed)
        return gwh_set_error(GWH_ERR_NOMEM, "launch needs " + std::to_string(bytes) + " bytes of a fixed scratch region of " +
                                                std::to_string(cap));
    if (p) {
        hipError_t e = hipFree(p);
        (void)e;
        p = nullptr;
        cap = 0;
    }
    size_t nb = gwh_grown_capacity(bytes, 0);
    if (hipMalloc(&p, nb) != hipSuccess) {
        p = nullptr;
        return gwh_set_error(GWH_ERR_NOMEM, "hipMalloc of " + std::to_string(nb) + " bytes failed");
    }
    cap = nb;
    return GWH_OK;
}

void DevBuf::release()
{
    if (p && !fixed) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

/* Bumped (under the engine lock) whenever a cached key may stop naming its event: an event destroyed or evicted,
 * the cache mode or the registered arrays changed.  Per-thread lookups taken at one generation stay valid
 * while it holds. */
static std::atomic<uint64_t> g_event_gen{1};
static void bump_event_generation() { g_event_gen.fetch_add(1, std::memory_order_acq_rel); }

uint64_t gwh_hash_bytes(const void *p, size_t n, uint64_t seed)
{
    /* four independent multiply-xorshift chains over interleaved words (the single chain's multiply latency
       bounded it to ~3 GB/s; content-keyed single calls hash every array on every call), folded at the end */
    const unsigned char *b = (const unsigned char *)p;
    uint64_t h[4] = {seed ^ 0x9e3779b97f4a7c15ULL, seed ^ 0xc2b2ae3d27d4eb4fULL, seed ^ 0x165667b19e3779f9ULL,
                     seed ^ 0x27d4eb2f165667c5ULL};
    size_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int k = 0; k < 4; k++) {
            uint64_t w;
            memcpy(&w, b + i + 8 * k, 8);
            h[k] ^= w;
            h[k] *= 0xff51afd7ed558ccdULL;
            h[k] ^= h[k] >> 29;
        }
    uint64_t r = h[0];
    for (int k = 1; k < 4; k++) {
        r ^= h[k] + 0x9e3779b97f4a7c15ULL + (r << 6) + (r >> 2);
        r *= 0xff51afd7ed558ccdULL;
    }
    for (; i < n; i++) {
        r ^= b[i];
        r *= 0x100000001b3ULL;
    }
    r ^= r >> 31;
    return r;
}

/* natural cubic spline second derivatives / 2 (the gsl_interp_cspline system), host side */
static void spline_c(const double *x, const double *y, int n, double *c)
{
    std::vector<double> dg(n), rh(n), of(n);
    c[0] = c[n - 1] = 0;
    int m = n - 2;
    for (int i = 0; i < m; i++) {
        double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
        dg[i] = 2.0 * (h0 + h1);
        of[i] = h1;
        rh[i] = 3.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
    }
    for (int i = 1; i < m; i++) {
        double w = of[i - 1] / dg[i - 1];
        dg[i] -= w * of[i - 1];
        rh[i] -= w * rh[i - 1];
    }
    c[m] = rh[m - 1] / dg[m - 1];
    for (int i = m - 2; i >= 0; i--) c[i + 1] = (rh[i] - of[i] * c[i + 2]) / dg[i];
}

static const double k_geometric_factor[GWH_NDET_TABLE] = {1., 1., 1., 1., 1., 1., 0.86602540378, 0.86602540378,
                                                          0.86602540378};

/* Scratch of the allocation-free likelihood entry point (gwh_loglike_batch_device_noalloc): instance
 * records, partial sums and tile counts of one launch, GWH_NOALLOC_SCRATCH_BYTES on every device the library
 * runs on, held until the process ends.  Host code uses the address hipGetSymbolAddress returns for it
 * (Engine::noalloc_scratch), never the name. */
static __device__ __align__(256) unsigned char scratch[GWH_NOALLOC_SCRATCH_BYTES];

int gwh_ensure_init()
{
    Engine &E = gwh_engine();
    if (E.ready) return GWH_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return gwh_set_error(GWH_ERR_HIP, "no HIP device visible: the engine has no CPU fallback");
    if (E.device >= ndev) E.device = 0;
    GWH_HIP(hipSetDevice(E.device));
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, E.device) == hipSuccess && cu > 0)
        E.num_cu = cu;
    GWH_HIP(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
    E.fisher.stream = E.stream;
    GWH_HIP(hipEventCreate(&E.ev0));
    GWH_HIP(hipEventCreate(&E.ev1));
    DetGeom geo[GWH_NDET_TABLE];
    for (int d = 0; d < GWH_NDET_TABLE; d++) {
        for (int i = 0; i < 9; i++) geo[d].tensor[i] = gwh_det_tensor[d][i];
        for (int i = 0; i < 3; i++) geo[d].loc[i] = gwh_det_location[d][i];
        geo[d].gfac = k_geometric_factor[d];
    }
    GWH_HIP(hipMalloc(&E.d_geo, sizeof(geo)));
    GWH_HIP(hipMemcpy(E.d_geo, geo, sizeof(geo), hipMemcpyHostToDevice));
    const int n = GWH_QNM_N;
    std::vector<double> host(5 * n);
    memcpy(&host[0], gwh_qnm_spin, n * sizeof(double));
    memcpy(&host[n], gwh_qnm_fring, n * sizeof(double));
    memcpy(&host[2 * n], gwh_qnm_fdamp, n * sizeof(double));
    spline_c(gwh_qnm_spin, gwh_qnm_fring, n, &host[3 * n]);
    spline_c(gwh_qnm_spin, gwh_qnm_fdamp, n, &host[4 * n]);
    GWH_HIP(hipMalloc(&E.d_qnm, host.size() * sizeof(double)));
    GWH_HIP(hipMemcpy(E.d_qnm, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
    E.qnm.spin = E.d_qnm;
    E.qnm.fring = E.d_qnm + n;
    E.qnm.fdamp = E.d_qnm + 2 * n;
    E.qnm.c_ring = E.d_qnm + 3 * n;
    E.qnm.c_damp = E.d_qnm + 4 * n;
    E.qnm.n = n;
    E.qnm.x0 = gwh_qnm_spin[0];
    E.qnm.x1 = gwh_qnm_spin[n - 1];
    GWH_HIP(hipMalloc(&E.d_events, sizeof(EventDesc) * GWH_MAX_EVENTS));
    GWH_HIP(hipMemset(E.d_events, 0, sizeof(EventDesc) * GWH_MAX_EVENTS));
    E.events.assign(GWH_MAX_EVENTS, Event());
    GWH_HIP(hipGetSymbolAddress((void **)&E.noalloc_scratch, HIP_SYMBOL(scratch)));
    if (gwh_tiles_load<GWH_IMRPHENOMD, GWH_FP64>() || gwh_tiles_load<GWH_IMRPHENOMD, GWH_MIXED>() ||
        gwh_tiles_load<GWH_IMRPHENOMPV2, GWH_FP64>() || gwh_tiles_load<GWH_IMRPHENOMPV2, GWH_MIXED>())
        return gwh_set_error(GWH_ERR_HIP, "loading the likelihood kernels failed");
    E.ready = true;
    return GWH_OK;
}

/* ------------------------------------------------------------------------- */
/* event preparation                                                          */
/* ------------------------------------------------------------------------- */

/* the tabulated powers of every bin (five arrays of `stride` doubles) */
__global__ void gwh_bin_powers_kernel(const double *__restrict__ f, int N, double *__restrict__ out, size_t stride)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    BinPow b = gwh_bin_powers(f[k]);
    out[k] = b.lf;
    out[stride + k] = b.f13;
    out[2 * stride + k] = b.fm13;
    out[3 * stride + k] = b.fm76;
    out[4 * stride + k] = b.f34;
}

/* weights folded with 4/S: Gauss-Legendre (the caller's) or Simpson by index parity; log10_jac multiplies by
 * f ln 10 (a log10 frequency axis: the likelihood applies it to Gauss-Legendre weights only, mcmc_gw.cpp:821-856,
 * calculate_snr_internal and calculate_fisher_elements under either rule, waveform_util.cpp:489-491,
 * fisher.cpp:2726-2728) */
__global__ void gwh_event_prep_kernel(const double *__restrict__ f, const double *__restrict__ psd,
                                      const double2 *__restrict__ data, const double *__restrict__ wts, int N, int D,
                                      int method, int log10_jac, double dfac, double *__restrict__ wS,
                                      double2 *__restrict__ dwS)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int d = blockIdx.y;
    if (k >= N || d >= D) return;
    double w;
    if (method == GWH_GAUSSLEG)
        w = 4.0 * wts[k];
    else
        w = 4.0 * dfac * simpson_w(k, N);
    if (log10_jac) w *= f[k] * 2.302585092994045684017991454684364208;
    double s = w / psd[(size_t)d * N + k];
    wS[(size_t)d * N + k] = s;
    double2 x = data[(size_t)d * N + k];
    dwS[(size_t)d * N + k] = make_double2(x.x * s, x.y * s);
}


/* element offsets of the frequency-minimum levels of an N-bin grid (level L: ceil(N / (64 * 2^L)) chunks) and
 * their total */
struct FminLevels {
    int off[GWH_FMIN_LEVELS];
};
static int fmin_levels(int N, FminLevels &L)
{
    int total = 0;
    for (int l = 0; l < GWH_FMIN_LEVELS; l++) {
        L.off[l] = total;
        total += (N + (64 << l) - 1) / (64 << l);
    }
    return total;
}

/* Smallest frequency of every chunk of 64 * 2^L bins, L = 0..7: a tile of lanes x bins-per-lane bins is one chunk of
 * its level, so a work item reads one value to know whether the tile starts above the 0.2/M cutoff.  One block of
 * 128 lanes per 8192 bins: each lane takes the minimum of its 64 bins (level 0), then pairs are combined level by
 * level in LDS. */
__global__ __launch_bounds__(128) void gwh_fmin_levels_kernel(const double *__restrict__ f, int N,
                                                              double *__restrict__ out, FminLevels L)
{
    __shared__ double s[2][128];
    const int t = threadIdx.x;
    const int c0 = blockIdx.x * 128 + t; /* level-0 chunk */
    double m = __longlong_as_double(0x7ff0000000000000LL);
    for (int i = 0; i < 64; i++) {
        const long long k = (long long)c0 * 64 + i;
        if (k < N) m = fmin(m, f[k]);
    }
    if (c0 < (N + 63) / 64) out[L.off[0] + c0] = m;
    s[0][t] = m;
    __syncthreads();
    int src = 0;
    for (int l = 1; l < GWH_FMIN_LEVELS; l++) {
        const int width = 128 >> l;
        if (t < width) {
            const double v = fmin(s[src][2 * t], s[src][2 * t + 1]);
            s[src ^ 1][t] = v;
            const int c = blockIdx.x * width + t;
            if (c < (N + (64 << l) - 1) / (64 << l)) out[L.off[l] + c] = v;
        }
        __syncthreads();
        src ^= 1;
    }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int gwh_event_create_ex(const double *f, int N, int ndet, const int *dets, const double *data, const double *psd,
                        int method, const double *weights, int log10F, int df_mode, uint64_t key, int *event)
{
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    Engine &E = gwh_engine();
    if (!f || !data || !psd || !dets || !event) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    if (N < 2) return gwh_set_error(GWH_ERR_INVALID, "need at least 2 frequency bins");
    if (ndet < 1 || ndet > GWH_MAXDET) return gwh_set_error(GWH_ERR_INVALID, "1 to 4 detectors per event");
    for (int d = 0; d < ndet; d++)
        if (dets[d] < 0 || dets[d] >= GWH_NDET_TABLE)
            return gwh_set_error(GWH_ERR_UNSUPPORTED, "unsupported detector id " + std::to_string(dets[d]));
    if (method == GWH_GAUSSLEG && !weights) return gwh_set_error(GWH_ERR_INVALID, "GAUSSLEG needs weights");
    int slot = -1;
    /* the last slot is reserved for the single-call waveform/response descriptor */
    for (int i = 0; i < GWH_MAX_EVENTS - 1; i++)
        if (!E.events[i].live) {
            slot = i;
            break;
        }
    if (slot < 0) return gwh_set_error(GWH_ERR_NOMEM, "event table full");
    const int nchunk = (N + GWH_TPB - 1) / GWH_TPB;
    FminLevels levels;
    const int nfmin = fmin_levels(N, levels);
    size_t Nd = (size_t)N * ndet;
    size_t off_f = 0, off_wS = align256(off_f + N * 8), off_dwS = align256(off_wS + Nd * 8),
           off_cm = align256(off_dwS + Nd * 16), off_pw = align256(off_cm + (size_t)nfmin * 8),
           total = align256(off_pw + 5 * align256((size_t)N * 8));
    const size_t pw_stride = align256((size_t)N * 8);
    char *blk = nullptr;
    if (hipMalloc((void **)&blk, total) != hipSuccess) return gwh_set_error(GWH_ERR_NOMEM, "event allocation failed");
    /* staging: psd, data, weights */
    int rc = E.psd.ensure(Nd * 8);
    if (rc == GWH_OK) rc = E.data.ensure(Nd * 16);
    if (rc == GWH_OK && weights) rc = E.wts.ensure((size_t)N * 8);
    if (rc != GWH_OK) {
        (void)hipFree(blk);
        return rc;
    }
    GWH_HIP(hipMemcpyAsync(blk + off_f, f, N * 8, hipMemcpyHostToDevice, E.stream));
    GWH_HIP(hipMemcpyAsync(E.psd.p, psd, Nd * 8, hipMemcpyHostToDevice, E.stream));
    GWH_HIP(hipMemcpyAsync(E.data.p, data, Nd * 16, hipMemcpyHostToDevice, E.stream));
    if (weights) GWH_HIP(hipMemcpyAsync(E.wts.p, weights, (size_t)N * 8, hipMemcpyHostToDevice, E.stream));
    double dfac = (df_mode == 0) ? (f[N / 2] - f[N / 2 - 1]) / 3.0 : (f[1] - f[0]) / 3.0;
    dim3 grid((N + GWH_TPB - 1) / GWH_TPB, ndet);
    const int log10_jac = log10F && (method == GWH_GAUSSLEG || df_mode == 1);
    hipLaunchKernelGGL(gwh_event_prep_kernel, grid, dim3(GWH_TPB), 0, E.stream, (const double *)(blk + off_f),
                       E.psd.as<const double>(), E.data.as<const double2>(), E.wts.as<const double>(), N, ndet, method,
                       log10_jac, dfac, (double *)(blk + off_wS), (double2 *)(blk + off_dwS));
    GWH_HIP(hipGetLastError());
    hipLaunchKernelGGL(gwh_fmin_levels_kernel, dim3((N + 8191) / 8192), dim3(128), 0, E.stream,
                       (const double *)(blk + off_f), N, (double *)(blk + off_cm), levels);
    GWH_HIP(hipGetLastError());
    hipLaunchKernelGGL(gwh_bin_powers_kernel, dim3(nchunk), dim3(GWH_TPB), 0, E.stream,
                       (const double *)(blk + off_f), N, (double *)(blk + off_pw), pw_stride / 8);
    GWH_HIP(hipGetLastError());
    Event &ev = E.events[slot];
    ev.live = true;
    ev.N = N;
    ev.D = ndet;
    ev.block = blk;
    ev.bytes = total;
    ev.key = key;
    ev.method = method;
    ev.log10F = log10F;
    ev.last_use = ++E.use_clock;
    /* the stored weights carry the step of df_mode (df_like below); the Fisher rescales to f[1]-f[0] */
    double df_l = f[N / 2] - f[N / 2 - 1], df_f = f[1] - f[0];
    ev.df_like = (df_mode == 0) ? df_l : df_f;
    ev.df_fisher = (method == GWH_GAUSSLEG) ? ev.df_like : df_f;
    EventDesc &ds = ev.desc;
    memset(&ds, 0, sizeof(ds));
    ds.f = (const double *)(blk + off_f);
    ds.wS = (const double *)(blk + off_wS);
    ds.dwS = (const double2 *)(blk + off_dwS);
    ds.tile_fmin = (const double *)(blk + off_cm);
    for (int l = 0; l < GWH_FMIN_LEVELS; l++) ds.fmin_off[l] = levels.off[l];
    ds.lgf = (const double *)(blk + off_pw);
    ds.f13 = (const double *)(blk + off_pw + pw_stride);
    ds.fm13 = (const double *)(blk + off_pw + 2 * pw_stride);
    ds.fm76 = (const double *)(blk + off_pw + 3 * pw_stride);
    ds.f34 = (const double *)(blk + off_pw + 4 * pw_stride);
    ds.N = N;
    ds.D = ndet;
    ds.bpl = gwh_choose_bpl(N);
    ds.lanes = gwh_choose_lanes(N);
    /* Gauss-Legendre weights carry no step: nothing to rescale */
    ds.snr_scale = (method == GWH_GAUSSLEG) ? 1.0 : df_f / ev.df_like;
    ds.sorted = 1;
    for (int k = 1; k < N; k++)
        if (!(f[k] >= f[k - 1])) {
            ds.sorted = 0;
            break;
        }
    for (int d = 0; d < GWH_MAXDET; d++) {
        ds.det[d] = (d < ndet) ? dets[d] : dets[0];
        ev.det[d] = ds.det[d];
    }
    GWH_HIP(hipMemcpyAsync(E.d_events + slot, &ds, sizeof(ds), hipMemcpyHostToDevice, E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    *event = slot;
    return GWH_OK;
}

/* ------------------------------------------------------------------------- */
/* setup                                                                      */
/* ------------------------------------------------------------------------- */

/* 64 instances per block, four waves (roles) per instance group: each lane computes one part of its
 * instance's record, the parts meet in LDS (see gwh_model.h for the split).  256 threads per block leave
 * each lane the whole register file of its SIMD, so the FP64 working set needs no scratch memory.  Every
 * wave reaches both barriers; lanes without an instance only skip the work.  MOD: the ppE / gIMR / theory
 * modifications are compiled in (the GR models run a kernel without them, which keeps its registers and
 * code to what GR needs). */
template <int MODEL, bool MOD>
__global__ __launch_bounds__(256) void gwh_setup_kernel(const gwh_params *__restrict__ params,
                                                        const int *__restrict__ inst_event, int event0, int n,
                                                        QnmTable qnm,
                                                        const DetGeom *__restrict__ geo,
                                                        const EventDesc *__restrict__ events,
                                                        InstCoef *__restrict__ coef, int *__restrict__ active,
                                                        const int *__restrict__ d_count, const int *__restrict__ d_mask,
                                                        int ppe_variant)
{
    const int ppe = MOD ? ppe_variant : 0;
    __shared__ SetupShared sh[64];
    constexpr bool pv2 = (MODEL == GWH_IMRPHENOMPV2);
    const int lane = threadIdx.x & 63;
    const int role = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = blockIdx.x * 64 + lane;
    bool live = i < n;
    if (live && ((d_count && i >= *d_count) || (d_mask && !d_mask[i]))) {
        if (active && role == 2) active[i] = -1; /* not evaluated: no tiles, NaN result */
        live = false;
    }
    SetupShared &X = sh[lane];
    InstCoef *C = coef + (live ? i : 0);
    gwh_params p;
    SetupState S;
    int ev = event0;
    double eul_a = 0, eul_e = 0; /* Pv2: Euler angles at f_ref (role 3) */

    /* ---- phase 1 ---- */
    if (live) {
        if (inst_event) ev = inst_event[i];
        p = params[i];
        gwh_setup_base(p, ppe, S);
        if (role == 0) {
            const double chip = pv2 ? gwh_setup_chip(p, S) : -1;
            gwh_setup_ringdown(p, pv2, chip, &gwh_phenomd_fit[0][0], qnm, S);
            gwh_setup_amp_mr(S, S.f3, &X.v3, &X.d3);
            SetupState &T = X.S; /* base, ringdown fits and frequencies */
            T.m1 = S.m1;
            T.m2 = S.m2;
            T.M = S.M;
            T.eta = S.eta;
            T.chi1 = S.chi1;
            T.chi2 = S.chi2;
            T.chi_s = S.chi_s;
            T.chi_a = S.chi_a;
            T.delta = S.delta;
            T.chi_pn = S.chi_pn;
            T.chi_eff = S.chi_eff;
            T.fRD = S.fRD;
            T.fdamp = S.fdamp;
            T.f1 = S.f1;
            T.f3 = S.f3;
            T.f1p = S.f1p;
            T.f2p = S.f2p;
            for (int k = 0; k < 3; k++) T.g[k] = S.g[k];
            T.ppe_n = S.ppe_n;
            T.ppe_imr = S.ppe_imr;
            T.gimr = S.gimr;
            for (int k = 0; k < 4; k++) T.gimr_neg[k] = S.gimr_neg[k];
            C->M = S.M;
            C->eta = S.eta;
            C->fcut = .2 / S.M;
            C->f1 = S.f1;
            C->f3 = S.f3;
            C->f1p = S.f1p;
            C->f2p = S.f2p;
            C->fRD = S.fRD;
            C->fdamp = S.fdamp;
            C->inv_fdamp = 1.0 / S.fdamp;
            C->mr_k = S.g[0] * S.g[2] * S.fdamp / S.M;
            C->mr_g2 = S.g[1] / (S.g[2] * S.fdamp);
            C->mr_w2 = (S.g[2] * S.fdamp) * (S.g[2] * S.fdamp);
        } else if (role == 1) {
            gwh_setup_pn(p, S);
            const double xfit = S.chi_pn - 1;
            for (int k = 0; k < 3; k++) S.rho[k] = gwh_fit(&gwh_phenomd_fit[0][0], k, S.eta, xfit);
            gwh_setup_amp_ins(S, 0.014 / S.M, &X.v1, &X.d1);
            for (int k = 0; k < 3; k++) X.S.rho[k] = S.rho[k];
            for (int k = 0; k < 7; k++) X.S.pa[k] = S.pa[k];
            for (int k = 0; k < 12; k++) X.S.pp[k] = S.pp[k];
            double pi13 = cbrt(GWH_PI), pw = 1.0;
            for (int k = 0; k < 7; k++) {
                C->ca[k] = S.pa[k] * pw;
                pw *= pi13;
            }
            for (int k = 0; k < 3; k++) C->rho[k] = S.rho[k];
            C->K = 3. / (128. * S.eta);
            for (int k = 0; k < 8; k++) C->cp[k] = S.pp[k];
            C->c8 = S.pp[8];
            C->c9 = S.pp[9];
            C->c10 = S.pp[10];
            C->c11 = S.pp[11];
            C->gimr = S.gimr;
            for (int k = 0; k < 4; k++) C->gimr_neg[k] = S.gimr_neg[k];
        } else if (role == 2) {
            double zeta = 0;
            if (pv2) {
                const Pv2Frame F = gwh_setup_frame(p, S, gwh_setup_chip(p, S));
                X.alpha0 = F.alpha0;
                X.phi_aligned = F.phi_aligned;
                zeta = F.zeta;
                double ct = cos(F.thetaJN), st = sin(F.thetaJN);
                C->ylm[0] = sqrt(5.0 / (64.0 * GWH_PI)) * (1.0 - ct) * (1.0 - ct);
                C->ylm[1] = sqrt(5.0 / (16.0 * GWH_PI)) * st * (1.0 - ct);
                C->ylm[2] = sqrt(15.0 / (32.0 * GWH_PI)) * st * st;
                C->ylm[3] = sqrt(5.0 / (16.0 * GWH_PI)) * st * (1.0 + ct);
                C->ylm[4] = sqrt(5.0 / (64.0 * GWH_PI)) * (1.0 + ct) * (1.0 + ct);
            }
            const EventDesc &E = events[ev];
            gwh_setup_detectors(p, pv2, zeta, geo, E.det, E.D, *C);
        } else {
            X.alpha1_fit = gwh_setup_phase_fits(p, &gwh_phenomd_fit[0][0], S);
            {
                SetupState &T = X.S; /* phase and intermediate-amplitude fits */
                T.v2 = S.v2;
                for (int k = 0; k < 5; k++) T.sigma[k] = S.sigma[k];
                for (int k = 0; k < 4; k++) T.beta[k] = S.beta[k];
                for (int k = 0; k < 6; k++) T.alpha[k] = S.alpha[k];
            }
            C->sig[0] = S.sigma[1] / S.eta;
            C->sig[1] = 0.75 * S.sigma[2] / S.eta;
            C->sig[2] = 0.6 * S.sigma[3] / S.eta;
            C->sig[3] = 0.5 * S.sigma[4] / S.eta;
            gwh_setup_mass(p, ppe, S);
            X.S.Mc = S.Mc;
            X.S.A0 = S.A0;
            for (int k = 0; k < GWH_PPE_MAX; k++) {
                X.S.ppe_e[k] = S.ppe_e[k];
                X.S.ppe_c[k] = S.ppe_c[k];
            }
            const double M76 = pow(S.M, 7. / 6.);
            C->amp_pref = pv2 ? S.A0 * M76 / (2. * sqrt(5. / (64. * GWH_PI))) / 2. : S.A0 * M76;
            C->M13 = cbrt(S.M);
            C->Mm13 = 1.0 / C->M13;
            C->Mm76 = 1.0 / (S.M * sqrt(C->M13));
            C->M34 = sqrt(sqrt(S.M * S.M * S.M));
            C->lgM = log(S.M);
            C->lgPiM = log(GWH_PI * S.M);
            C->ppe_n = S.ppe_n;
            C->ppe_imr = S.ppe_imr;
            for (int k = 0; k < GWH_PPE_MAX; k++) {
                C->ppe_e[k] = S.ppe_e[k];
                C->ppe_c[k] = S.ppe_c[k];
            }
            if (pv2) { /* precession constants (IMRPhenomP.cpp:64-144, :1081-1211) */
                const double chip = gwh_setup_chip(p, S);
                const double q = S.m1 / S.m2;
                const double chil = (1.0 + q) / q * S.chi_eff;
                const double mm1 = q / (1 + q), mm2 = 1. / (1 + q);
                C->SP = chip * mm1 * mm1;
                C->SL = S.chi1 * mm1 * mm1 + S.chi2 * mm2 * mm2;
                C->L_c1 = 1.5 + S.eta / 6.0;
                C->L_c2 = 3.375 - (19.0 * S.eta) / 8. - S.eta * S.eta / 24.0;
                double a_c[5], e_c[5];
                gwh_euler_coeffs(q, chil, chip, a_c, e_c);
                for (int k = 0; k < 5; k++) {
                    C->a_c[k] = a_c[k];
                    C->e_c[k] = e_c[k];
                }
                double vr = cbrt(GWH_PI * S.M * p.f_ref);
                double lw = log(vr * vr * vr);
                eul_a = gwh_euler_eval(a_c, vr, lw);
                eul_e = gwh_euler_eval(e_c, vr, lw);
            }
        }
        /* the tile count (a dependent chain of loads) on a role with a short chain: the detectors role for
           IMRPhenomD, the ringdown role for IMRPhenomPv2, whose frame makes the detectors role the longest */
        if (role == (pv2 ? 0 : 2)) {
            if (active) {
                const EventDesc &E = events[ev];
                /* tiles to visit: all of them, or on an ascending grid those starting at or below 0.2/M */
                const double fcut = .2 / S.M;
                const int tile = E.lanes * E.bpl;
                int nt = (E.N + tile - 1) / tile;
                const double *tf = E.tile_fmin + E.fmin_off[gwh_tile_level(tile)]; /* tile minima */
                if (E.sorted) {
                    /* the first tile whose smallest frequency exceeds the cutoff; up to 16 tiles the count of
                     * tiles at or below it (same answer on an ascending grid), from loads issued together */
                    int lo = 0;
                    if (nt <= 0) {
                        lo = 0;
                    } else if (nt <= 16) {
                        double fm[16];
#pragma unroll
                        for (int t = 0; t < 16; t++) fm[t] = tf[t < nt ? t : nt - 1];
#pragma unroll
                        for (int t = 0; t < 16; t++) lo += (t < nt && !(fm[t] > fcut));
                    } else {
                        int hi = nt;
                        while (lo < hi) {
                            int mid = (lo + hi) >> 1;
                            if (tf[mid] > fcut)
                                hi = mid;
                            else
                                lo = mid + 1;
                        }
                    }
                    nt = (fcut == fcut) ? lo : nt; /* NaN cutoff: visit everything, NaN propagates */
                }
                active[i] = nt;
            }
        }
    }
    __syncthreads();

    /* ---- phase 2: connection pieces with beta0 = beta1 = alpha0 = alpha1 = 0 ---- */
    if (live) {
        S = X.S;
        S.beta[0] = S.beta[1] = 0;
        S.alpha[0] = S.alpha[1] = 0;
        if (role == 0) {
            X.p_ins = gwh_setup_phase_ins(S, S.f1p);
            X.l1 = log(S.M * S.f1p);
        } else if (role == 1) {
            X.d_ins = gwh_setup_dphase_ins(S, S.f1p);
            X.l2 = log(S.M * S.f2p);
        } else if (role == 2) {
            X.q_mr = gwh_setup_phase_mr(S, S.f2p);
            X.dq_mr = gwh_setup_dphase_mr(S, S.f2p);
            double dl[5];
            gwh_setup_amp_connection(S.M, S.f1, S.f3, X.v1, X.d1, S.v2, X.v3, X.d3, dl);
            for (int k = 0; k < 5; k++) C->delta[k] = dl[k];
        } else {
            if (!pv2) {
                if (p.shift_phase) {
                    const double fr = p.f_ref;
                    X.ref = fr < S.f1p ? gwh_setup_phase_ins(S, fr)
                                       : (fr > S.f2p ? gwh_setup_phase_mr(S, fr) : log(S.M * fr));
                }
                if (p.shift_time) X.dq_mr3 = gwh_setup_dphase_mr(S, S.f3);
            } else {
                C->alpha_add = X.alpha0 - eul_a;
                C->eps_add = -eul_e;
            }
        }
        if (pv2 && p.shift_time)
            for (int j = role; j < 10; j += 4) X.tq[j] = gwh_setup_phase_mr(S, gwh_setup_knot(S.fRD, j));
    }
    __syncthreads();

    /* ---- phase 3: the phase connection (IMRPhenomD.cpp:1614-1674) and the alignment ---- */
    if (live && role == 0) {
        const double eta = S.eta, M = S.M;
        S.beta[1] = (eta / M) * X.d_ins - (eta / M) * gwh_setup_dphase_int(S, S.f1p);
        S.beta[0] = eta * X.p_ins - eta * gwh_setup_phase_int_l(S, S.f1p, X.l1);
        S.alpha[1] = (eta / M) * gwh_setup_dphase_int(S, S.f2p) - (eta / M) * X.dq_mr;
        S.alpha[0] = eta * gwh_setup_phase_int_l(S, S.f2p, X.l2) -
                     eta * (X.q_mr + (1. / eta) * (S.alpha[1] * (M * S.f2p)));
        C->beta[0] = S.beta[0] / eta;
        C->beta[1] = S.beta[1] / eta;
        C->beta[2] = S.beta[2] / eta;
        C->beta[3] = S.beta[3] / (3. * eta);
        C->alpha[0] = S.alpha[0] / eta;
        C->alpha[1] = S.alpha[1] / eta;
        C->alpha[2] = S.alpha[2] / eta;
        C->alpha[3] = (4. / 3.) * S.alpha[3] / eta;
        C->alpha[4] = S.alpha[4] / eta;
        C->alpha[5] = S.alpha[5] * S.fRD;
        if (!pv2) {
            /* phase and time alignment (IMRPhenomD.cpp:441-466) */
            double phic, fref;
            if (p.shift_phase) {
                fref = p.f_ref;
                double b;
                if (fref < S.f1p)
                    b = X.ref;
                else if (fref > S.f2p)
                    b = X.ref + (1. / eta) * (S.alpha[0] + S.alpha[1] * (M * fref));
                else
                    b = gwh_setup_phase_int_l(S, fref, X.ref);
                phic = 2 * p.phiRef + b;
            } else {
                fref = 0;
                phic = p.phiRef;
            }
            double tc_shift = 0;
            if (p.shift_time)
                tc_shift = (X.dq_mr3 + (1. / eta) * (M * S.alpha[1])) + (-S.alpha[1] + X.alpha1_fit) * M / eta;
            C->tc_total = GWH_2PI * p.tc + tc_shift;
            C->phic = phic;
            C->f_ref = fref;
            C->t_corr_2pi = 0;
        } else {
            C->phic = 2 * X.phi_aligned;
            C->tc_total = GWH_2PI * p.tc;
            C->f_ref = p.f_ref;
            /* t_c alignment: derivative at fRD of a 10-knot natural spline of -Psi_MR (IMRPhenomP.cpp:550-606) */
            double t_corr = 0;
            if (p.shift_time) {
                double ph[10];
#pragma unroll
                for (int j = 0; j < 10; j++)
                    ph[j] = -(X.tq[j] + (1. / eta) * (S.alpha[0] + S.alpha[1] * (M * gwh_setup_knot(S.fRD, j))));
                t_corr = gwh_setup_tcorr(S.fRD, ph);
            }
            C->t_corr_2pi = GWH_2PI * t_corr;
        }
    }
}

static void launch_setup(int model, int n, hipStream_t st, const gwh_params *params, const int *inst_event,
                         int event0, InstCoef *coef, int *active, const int *d_count = nullptr,
                         const int *d_mask = nullptr)
{
    Engine &E = gwh_engine();
    dim3 g((n + 63) / 64), b(256);
    const int var = gwh_ppe_variant(model);
    auto k = gwh_is_pv2(model) ? (var ? gwh_setup_kernel<GWH_IMRPHENOMPV2, true> : gwh_setup_kernel<GWH_IMRPHENOMPV2, false>)
                               : (var ? gwh_setup_kernel<GWH_IMRPHENOMD, true> : gwh_setup_kernel<GWH_IMRPHENOMD, false>);
    hipLaunchKernelGGL(k, g, b, 0, st, params, inst_event, event0, n, E.qnm, (const DetGeom *)E.d_geo,
                       (const EventDesc *)E.d_events, coef, active, d_count, d_mask, var);
}

int gwh_setup_launch(const int *d_events, int events0, int model, const gwh_params *d_params, int n, InstCoef *d_coef,
                     hipStream_t stream)
{
    Engine &E = gwh_engine();
    launch_setup(model, n, stream, d_params, d_events, events0, d_coef, nullptr);
    GWH_HIP(hipGetLastError());
    return GWH_OK;
}

/* Work list of one tiling class (single block): instance i contributes active[i] items if its event is tiled
 * with `bpl` bins per lane and `lanes` lanes per tile; items are ordered by instance then tile.  At most `cap` are
 * written. */
__global__ __launch_bounds__(1024) void gwh_worklist_kernel(WorkMap W, const EventDesc *__restrict__ events,
                                                            const int *__restrict__ active, int bpl, int lanes,
                                                            int2 *__restrict__ work, int *__restrict__ count, int cap)
{
    __shared__ int part[1024];
    const int t = threadIdx.x, n = W.n;
    const int chunk = (n + 1023) / 1024;
    const int i0 = min(n, t * chunk), i1 = min(n, i0 + chunk);
    int local = 0;
    for (int i = i0; i < i1; i++) {
        int ev = W.inst_event ? W.inst_event[i] : W.event0;
        if (events[ev].bpl == bpl && events[ev].lanes == lanes) local += max(active[i], 0);
    }
    part[t] = local;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) { /* inclusive scan */
        int v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int off = part[t] - local;
    for (int i = i0; i < i1; i++) {
        int ev = W.inst_event ? W.inst_event[i] : W.event0;
        if (events[ev].bpl != bpl || events[ev].lanes != lanes) continue;
        for (int k = 0; k < active[i]; k++, off++)
            if (off < cap) work[off] = make_int2(i, k);
    }
    if (t == 1023) *count = min(part[1023], cap);
}

/* One wave per instance: lane l sums tiles l, l + 64, ... in order, then a fixed xor-shuffle tree; the
 * order depends only on the instance's own tile count, so results do not depend on the batch. */
__global__ __launch_bounds__(GWH_TPB) void gwh_finalize_kernel(WorkMap W, const EventDesc *__restrict__ events,
                                                               const InstCoef *__restrict__ coefs,
                                                               const double *__restrict__ partial,
                                                               const int *__restrict__ active, int model, int what,
                                                               double *__restrict__ logl, const int *__restrict__ count)
{
    /* The instance index is uniform over the wave (scalar loads).  Each lane's first tile record (every tile slot
       lies inside the buffer) is requested first, then the loads that do not depend on each other (tile count,
       event id, the instance's detector constants); slots at or past the tile count are added as zeros, so the
       sums are those of the tiles in order, without a round trip for the count before the partial sums. */
    const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * (GWH_TPB / 64) + (threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    if (i >= W.n || (count && i >= *count)) return; /* whole waves leave together */
    const size_t t0 = (size_t)i * W.tiles0;
    double v0[GWH_PARTIAL_STRIDE];
    const bool has0 = lane < W.tiles0; /* lanes without a tile slot read nothing */
#pragma unroll
    for (int j = 0; j < GWH_PARTIAL_STRIDE; j++) v0[j] = 0.0;
    if (has0) {
        const double *p = partial + (t0 + lane) * GWH_PARTIAL_STRIDE;
#pragma unroll
        for (int j = 0; j < GWH_PARTIAL_STRIDE; j++) v0[j] = p[j];
    }
    const int nt = active[i];
    const int ev = W.inst_event ? W.inst_event[i] : W.event0;
    const InstCoef &C = coefs[i];
    double cdet[GWH_MAXDET][3];
#pragma unroll
    for (int d = 0; d < GWH_MAXDET; d++)
        for (int q = 0; q < 3; q++) cdet[d][q] = C.det[d][q];
    const EventDesc &E = events[ev];
    if (nt < 0) { /* instance not evaluated (count or mask of the launch) */
        if (lane == 0) logl[i] = __longlong_as_double(0x7ff8000000000000ll);
        return;
    }
    /* every slot of a tile's partial record is summed (slots past 3 D are unused), so the loads need not wait
       for the event's detector count */
    double s[GWH_PARTIAL_STRIDE];
    const bool use0 = has0 && lane < nt;
#pragma unroll
    for (int j = 0; j < GWH_PARTIAL_STRIDE; j++) s[j] = use0 ? v0[j] : 0.0;
    for (int t = lane + 64; t < nt; t += 64) {
        const double *p = partial + (t0 + t) * GWH_PARTIAL_STRIDE;
#pragma unroll
        for (int j = 0; j < GWH_PARTIAL_STRIDE; j++) s[j] += p[j];
    }
    const int D = E.D, nv = 3 * D;
#pragma unroll
    for (int j = 0; j < GWH_PARTIAL_STRIDE; j++)
        if (j < nv) s[j] = wave_sum(s[j]);
    if (lane != 0) return;
    double ll = 0;
#pragma unroll
    for (int d = 0; d < GWH_MAXDET; d++) {
        if (d >= D) break;
        double hh, dh;
        if (!gwh_is_pv2(model)) {
            double ha = s[3 * d], zr = s[3 * d + 1], zi = s[3 * d + 2];
            hh = cdet[d][2] * ha;
            dh = cdet[d][0] * zr + cdet[d][1] * zi;
        } else {
            hh = s[3 * d];
            dh = s[3 * d + 1];
        }
        ll += (what == 0) ? -0.5 * (hh - 2 * dh) : hh;
    }
    /* what 1: network optimal SNR sqrt(sum_d (h_d|h_d)) with the SNR's Simpson step */
    logl[i] = (what == 0) ? ll : sqrt(ll * E.snr_scale);
}

int gwh_choose_bpl(int N)
{
    Engine &E = gwh_engine();
    if (E.bpl > 0) return E.bpl;
    int lg = 0;
    while ((2 << lg) <= N && lg < 31) lg++;
    if (E.tune[lg] > 0) return E.tune[lg];
    /* 8 bins per lane amortise the per-tile reduction best (measured on the 16225-bin network grid:
     * 1/2/4/8 -> 92/96/77/72 us per 512-instance launch); short grids keep at least two tiles */
    if (N >= 2 * GWH_TPB * 8) return 8;
    if (N >= 2 * GWH_TPB * 4) return 4;
    if (N >= 2 * GWH_TPB * 2) return 2;
    return 1;
}

/* lanes per tile: the engine-wide override, else the autotuned choice of the grid's size bucket, else 256 */
int gwh_choose_lanes(int N)
{
    Engine &E = gwh_engine();
    if (E.lanes > 0) return E.lanes;
    int lg = 0;
    while ((2 << lg) <= N && lg < 31) lg++;
    if (E.tune_lanes[lg] > 0) return E.tune_lanes[lg];
    return GWH_TPB;
}

/* tiling classes: bins per lane {1, 2, 4, 8} x lanes per tile {64, 128, 256, 512, 1024} */
static const int kClassBpl[4] = {1, 2, 4, 8}, kClassLanes[5] = {64, 128, 256, 512, 1024};
static int gwh_class_of(const EventDesc &d)
{
    const int b = d.bpl == 1 ? 0 : d.bpl == 2 ? 1 : d.bpl == 4 ? 2 : 3;
    const int l = d.lanes == 64 ? 0 : d.lanes == 128 ? 1 : d.lanes == 256 ? 2 : d.lanes == 512 ? 3 : 4;
    return 5 * b + l;
}

static void launch_tiles(int model, int precision, int cls, long long items, const int2 *work, hipStream_t st,
                         const WorkMap &W, const EventDesc *ev, const InstCoef *c, double *part, const int *act,
                         const int *count)
{
    const int bpl = kClassBpl[cls / 5], lanes = kClassLanes[cls % 5];
    if (!gwh_is_pv2(model)) {
        if (precision == GWH_FP64)
            gwh_launch_tiles<GWH_IMRPHENOMD, GWH_FP64>(bpl, lanes, items, work, st, W, ev, c, part, act, count);
        else
            gwh_launch_tiles<GWH_IMRPHENOMD, GWH_MIXED>(bpl, lanes, items, work, st, W, ev, c, part, act, count);
    } else {
        if (precision == GWH_FP64)
            gwh_launch_tiles<GWH_IMRPHENOMPV2, GWH_FP64>(bpl, lanes, items, work, st, W, ev, c, part, act, count);
        else
            gwh_launch_tiles<GWH_IMRPHENOMPV2, GWH_MIXED>(bpl, lanes, items, work, st, W, ev, c, part, act, count);
    }
}

int gwh_loglike_launch(const int *d_events, int events0, int model, int precision, const gwh_params *d_params, int n,
                       double *d_logl, hipStream_t stream, bool timed, double *kernel_ms, const int *h_events,
                       int what, Scratch *scratch, const LaunchExtra *extra)
{
    Engine &E = gwh_engine();
    static const LaunchExtra kNone;
    const LaunchExtra &X = extra ? *extra : kNone;
    DevBuf &s_coef = scratch ? scratch->coef : E.coef, &s_partial = scratch ? scratch->partial : E.partial,
           &s_active = scratch ? scratch->active : E.active, &s_work = scratch ? scratch->work : E.work,
           &s_wcount = scratch ? scratch->wcount : E.wcount;
    if (n <= 0) return GWH_OK;
    if (!gwh_model_valid(model))
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "model must be IMRPhenomD, IMRPhenomPv2 or a ppE IMRPhenomD");
    if (precision != GWH_FP64 && precision != GWH_MIXED) return gwh_set_error(GWH_ERR_INVALID, "bad precision");
    if (events0 < 0 || events0 >= GWH_MAX_EVENTS || (!d_events && !E.events[events0].live))
        return gwh_set_error(GWH_ERR_INVALID, "unknown event " + std::to_string(events0));
    if (n > 65535) return gwh_set_error(GWH_ERR_INVALID, "at most 65535 instances per launch");
    /* each event carries its own tiling (bins per lane x lanes per tile); one launch per tiling class present
     * (instances of other classes leave at once), so an instance's partial sums never depend on the batch */
    int tiles_b[GWH_NCLASS] = {0};
    auto note = [&](const Event &ev) {
        const int t = gwh_tiles_of(ev.desc), sl = gwh_class_of(ev.desc);
        if (t > tiles_b[sl]) tiles_b[sl] = t;
    };
    /* multi-event batch whose instance -> event map is known on the host: flattened work lists with
     * per-class upper bounds (sum over instances of the event's tile count) */
    const bool wl = d_events && h_events;
    long long ub[GWH_NCLASS] = {0};
    if (wl) {
        for (int i = 0; i < n; i++) {
            const Event &ev = E.events[h_events[i]];
            note(ev);
            ub[gwh_class_of(ev.desc)] += gwh_tiles_of(ev.desc);
        }
        for (int c = 0; c < GWH_NCLASS; c++)
            if (ub[c] > (1ll << 30)) return gwh_set_error(GWH_ERR_INVALID, "work list too large");
    } else if (d_events) {
        for (auto &ev : E.events)
            if (ev.live) note(ev);
    } else {
        note(E.events[events0]);
    }
    int maxtiles = 0;
    for (int b = 0; b < GWH_NCLASS; b++) maxtiles = std::max(maxtiles, tiles_b[b]);
    if ((long long)n * maxtiles > INT_MAX) return gwh_set_error(GWH_ERR_INVALID, "too many (instance, tile) items");
    int rc;
    if ((rc = s_coef.ensure((size_t)n * sizeof(InstCoef))) != GWH_OK) return rc;
    if ((rc = s_partial.ensure((size_t)n * maxtiles * GWH_PARTIAL_STRIDE * sizeof(double))) != GWH_OK) return rc;
    if ((rc = s_active.ensure((size_t)n * sizeof(int))) != GWH_OK) return rc;
    if (wl) {
        long long total = 0;
        for (int c = 0; c < GWH_NCLASS; c++) total += ub[c];
        if ((rc = s_work.ensure((size_t)total * sizeof(int2))) != GWH_OK) return rc;
        if ((rc = s_wcount.ensure(GWH_NCLASS * sizeof(int))) != GWH_OK) return rc;
    }
    if (X.reserve_only) return GWH_OK;
    WorkMap W;
    W.inst_event = d_events;
    W.event0 = events0;
    W.n = n;
    W.tiles0 = maxtiles;
    launch_setup(model, n, stream, d_params, d_events, events0, s_coef.as<InstCoef>(), s_active.as<int>(), X.d_count,
                 X.d_mask);
    GWH_HIP(hipGetLastError());
    if (X.capturing) timed = false;
    if (timed) GWH_HIP(hipEventRecord(E.ev0, stream));
    const bool live = !X.capturing && E.kt_on && (E.kt_seq++ % E.kt_every) == 0 &&
                      2 * (size_t)E.kt_used + 1 < E.kt_ev.size();
    if (live) GWH_HIP(hipEventRecord(E.kt_ev[2 * E.kt_used], stream));
    const EventDesc *evp = (const EventDesc *)E.d_events;
    const InstCoef *cp = s_coef.as<const InstCoef>();
    double *pp = s_partial.as<double>();
    const int *ap = s_active.as<const int>();
    if (wl) {
        long long off = 0;
        for (int sl = 0; sl < GWH_NCLASS; sl++) {
            if (!ub[sl]) continue;
            int2 *wp = s_work.as<int2>() + off;
            int *cnt = s_wcount.as<int>() + sl;
            hipLaunchKernelGGL(gwh_worklist_kernel, dim3(1), dim3(1024), 0, stream, W, evp, ap, kClassBpl[sl / 5],
                               kClassLanes[sl % 5], wp, cnt, (int)ub[sl]);
            launch_tiles(model, precision, sl, ub[sl], wp, stream, W, evp, cp, pp, ap, cnt);
            off += ub[sl];
        }
    }
    for (int sl = 0; sl < GWH_NCLASS && !wl; sl++)
        if (tiles_b[sl]) launch_tiles(model, precision, sl, (long long)n * maxtiles, nullptr, stream, W, evp, cp, pp, ap,
                                      X.d_count);
    GWH_HIP(hipGetLastError());
    if (timed) GWH_HIP(hipEventRecord(E.ev1, stream));
    if (live) GWH_HIP(hipEventRecord(E.kt_ev[2 * E.kt_used++ + 1], stream));
    hipLaunchKernelGGL(gwh_finalize_kernel, dim3((n + GWH_TPB / 64 - 1) / (GWH_TPB / 64)), dim3(GWH_TPB), 0, stream,
                       W, evp, cp,
                       (const double *)pp, ap, model, what, d_logl, X.d_count);
    GWH_HIP(hipGetLastError());
    if (timed && kernel_ms) {
        GWH_HIP(hipEventSynchronize(E.ev1));
        float ms = 0;
        GWH_HIP(hipEventElapsedTime(&ms, E.ev0, E.ev1));
        *kernel_ms = ms;
    }
    return GWH_OK;
}

/* bins of every instance by IMRPhenomD region (for the algorithmic-work figure of benchmarks):
 * counts[GWH_NREGION_COUNTS i + 3 a + p] for the non-zero bins (f <= 0.2/M) in amplitude region a (0: f < f1,
 * 1: f1 <= f <= f3, 2: f > f3) and phase region p (0: f < f1_phase, 1: up to f2_phase, 2: above);
 * counts[GWH_NREGION_COUNTS i + 9] = bins above 0.2/M */
__global__ void gwh_region_count_kernel(const InstCoef *__restrict__ coefs, const EventDesc *__restrict__ events,
                                        int event0, int *__restrict__ counts)
{
    __shared__ int s[GWH_NREGION_COUNTS][GWH_TPB];
    const InstCoef &C = coefs[blockIdx.x];
    const EventDesc &E = events[event0];
    int c[GWH_NREGION_COUNTS];
    for (int j = 0; j < GWH_NREGION_COUNTS; j++) c[j] = 0;
    for (int k = threadIdx.x; k < E.N; k += GWH_TPB) {
        const double f = E.f[k];
        const int a = (f < C.f1) ? 0 : (f > C.f3) ? 2 : 1;
        const int p = (f < C.f1p) ? 0 : (f > C.f2p) ? 2 : 1;
        c[(f > C.fcut) ? 9 : 3 * a + p]++;
    }
    for (int j = 0; j < GWH_NREGION_COUNTS; j++) s[j][threadIdx.x] = c[j];
    __syncthreads();
    for (int w = GWH_TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int j = 0; j < GWH_NREGION_COUNTS; j++) s[j][threadIdx.x] += s[j][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < GWH_NREGION_COUNTS) counts[GWH_NREGION_COUNTS * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

/* ------------------------------------------------------------------------- */
/* waveform / response output kernels (single-call drop-in path)              */
/* ------------------------------------------------------------------------- */

/* mode 0: hplus/hcross; mode 1: detector response (detector 0 of the instance record) */
template <int MODEL>
__global__ void gwh_waveform_kernel(const double *__restrict__ f, int N, const InstCoef *__restrict__ coef, int mode,
                                    double2 *__restrict__ out1, double2 *__restrict__ out2, int det = 0)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const InstCoef &C = coef[0];
    double fk = f[k];
    double2 hp = make_double2(0, 0), hc = make_double2(0, 0);
    if (fk <= C.fcut) {
        const BinPow bp = gwh_bin_powers(fk);
        if (MODEL == GWH_IMRPHENOMD) {
            double A, phi, u, lx;
            gwh_phenomd_fp64(C, fk, bp, A, phi, u, lx);
            double s, c;
            gwh_sincos_table(phi, s, c);
            double hr = A * c, hi = -A * s; /* h = A e^{-i phi} */
            hp = make_double2(hr * C.wf_a, hi * C.wf_a);
            hc = make_double2(hi * C.wf_b, -hr * C.wf_b); /* -i h cos(iota) */
        } else {
            double A, Phi, hpr, hpi, hcr, hci;
            gwh_pv2_fp64(C, fk, bp, A, Phi, hpr, hpi, hcr, hci);
            double s, c;
            gwh_sincos_table(Phi, s, c);
            /* (A hpf) e^{-i Phi} */
            double pr = A * (hpr * c + hpi * s), pi = A * (hpi * c - hpr * s);
            double xr = A * (hcr * c + hci * s), xi = A * (hci * c - hcr * s);
            double c2z = C.wf_a, s2z = C.wf_b;
            hp = make_double2(c2z * pr + s2z * xr, c2z * pi + s2z * xi);
            hc = make_double2(c2z * xr - s2z * pr, c2z * xi - s2z * pi);
        }
    }
    if (mode == 0) {
        out1[k] = hp;
        out2[k] = hc;
    } else {
        /* F+ h+ + Fx hx; the record holds F+, Fx through the projection constants */
        if (MODEL == GWH_IMRPHENOMD) {
            /* response = h K with K = F+ (1+c^2)/2 - i Fx c : rebuild from h+ (= h k+) */
            double hr = (C.wf_a != 0) ? hp.x / C.wf_a : 0, hi = (C.wf_a != 0) ? hp.y / C.wf_a : 0;
            double Kr = C.det[det][0], Ki = C.det[det][1];
            hp = make_double2(hr * Kr - hi * Ki, hr * Ki + hi * Kr);
        } else {
            /* P h+0 + Q hx0 on the unrotated polarisations equals F+ h+ + Fx hx on the rotated ones */
            double c2z = C.wf_a, s2z = C.wf_b;
            double p0r = c2z * hp.x - s2z * hc.x, p0i = c2z * hp.y - s2z * hc.y;
            double x0r = s2z * hp.x + c2z * hc.x, x0i = s2z * hp.y + c2z * hc.y;
            hp = make_double2(C.det[det][0] * p0r + C.det[det][1] * x0r, C.det[det][0] * p0i + C.det[det][1] * x0i);
        }
        if (mode == 2) {
            /* time-of-flight phase e^{-2 pi i f dt} relative to the first detector (waveform_util.cpp:170-179) */
            double s, c;
            gwh_sincos_table(-C.det[det][3] * fk, s, c);
            hp = make_double2(hp.x * c - hp.y * s, hp.x * s + hp.y * c);
        }
        out1[k] = hp;
    }
}

/* plain weighted inner products of a given response (Log_Likelihood_internal with the response passed in) */
__global__ void gwh_inner_kernel(const double2 *__restrict__ dwS, const double2 *__restrict__ resp,
                                 const double *__restrict__ wS, int N, double *__restrict__ partial)
{
    __shared__ double lds[GWH_TPB / 64][2];
    double acc[2] = {0, 0};
    for (int k = blockIdx.x * GWH_TPB + threadIdx.x; k < N; k += gridDim.x * GWH_TPB) {
        double2 r = resp[k], d = dwS[k];
        acc[0] += (r.x * r.x + r.y * r.y) * wS[k];
        acc[1] += d.x * r.x + d.y * r.y;
    }
    block_reduce_store<2>(acc, partial + 2 * blockIdx.x, lds);
}

/* ------------------------------------------------------------------------- */
/* C ABI                                                                      */
/* ------------------------------------------------------------------------- */

extern "C" {

const char *gwh_last_error(void) { return g_err.c_str(); }

const char *gwh_version(void) { return "gwat-hip 0.1 (gfx950)"; }

int gwh_init(int device)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (!E.ready) E.device = device;
    return gwh_ensure_init();
}

int gwh_synchronize(void)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

int gwh_set_bins_per_lane(int bpl)
{
    if (!(bpl == 0 || bpl == 1 || bpl == 2 || bpl == 4 || bpl == 8))
        return gwh_set_error(GWH_ERR_INVALID, "bins per lane must be 0, 1, 2, 4 or 8");
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    E.bpl = bpl;
    /* re-tile the resident events (later results may differ in the last bits from earlier ones) */
    for (int i = 0; i < GWH_MAX_EVENTS; i++) {
        Event &ev = E.events[i];
        if (!ev.live || ev.pins > 0) continue; /* events held by a step graph keep their tiling */
        ev.desc.bpl = gwh_choose_bpl(ev.N);
        GWH_HIP(hipMemcpyAsync(E.d_events + i, &ev.desc, sizeof(EventDesc), hipMemcpyHostToDevice, E.stream));
    }
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

int gwh_set_tile_lanes(int lanes)
{
    if (!(lanes == 0 || lanes == 64 || lanes == 128 || lanes == 256 || lanes == 512 || lanes == 1024))
        return gwh_set_error(GWH_ERR_INVALID, "lanes per tile must be 0, 64, 128, 256, 512 or 1024");
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    E.lanes = lanes;
    /* re-tile the resident events (later results may differ in the last bits from earlier ones) */
    for (int i = 0; i < GWH_MAX_EVENTS; i++) {
        Event &ev = E.events[i];
        if (!ev.live || ev.pins > 0) continue; /* events held by a step graph keep their tiling */
        ev.desc.lanes = gwh_choose_lanes(ev.N);
        GWH_HIP(hipMemcpyAsync(E.d_events + i, &ev.desc, sizeof(EventDesc), hipMemcpyHostToDevice, E.stream));
    }
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

void *gwh_device_alloc(size_t bytes)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (gwh_ensure_init() != GWH_OK) return nullptr;
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        gwh_set_error(GWH_ERR_NOMEM, "hipMalloc failed");
        return nullptr;
    }
    return p;
}

int gwh_device_free(void *ptr)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    GWH_HIP(hipFree(ptr));
    return GWH_OK;
}

int gwh_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    GWH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

int gwh_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    GWH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

int gwh_event_create(const double *frequencies, int length, int ndet, const int *detectors, const double *data,
                     const double *psd, int *event)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    return gwh_event_create_ex(frequencies, length, ndet, detectors, data, psd, GWH_SIMPSONS, nullptr, 0, 0, 0, event);
}

/* drop the pointer-keyed cached events (their host arrays may change once no longer registered) */
static int drop_pointer_keyed_events(Engine &E)
{
    for (int i = 0; i < GWH_MAX_EVENTS - 1; i++) {
        Event &ev = E.events[i];
        if (!ev.live || ev.key == 0 || !ev.ptr_key || ev.pins > 0) continue;
        GWH_HIP(hipStreamSynchronize(E.stream));
        GWH_HIP(hipFree(ev.block));
        ev = Event();
        bump_event_generation();
    }
    return GWH_OK;
}

int gwh_set_single_call_cache(int mode)
{
    if (mode != GWH_CACHE_BY_CONTENT && mode != GWH_CACHE_BY_POINTER)
        return gwh_set_error(GWH_ERR_INVALID, "cache mode must be GWH_CACHE_BY_CONTENT or GWH_CACHE_BY_POINTER");
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (E.cache_by_pointer && mode == GWH_CACHE_BY_CONTENT && (st = drop_pointer_keyed_events(E)) != GWH_OK) return st;
    E.cache_by_pointer = (mode == GWH_CACHE_BY_POINTER);
    bump_event_generation();
    return GWH_OK;
}


int gwh_host_data_register(int num_detectors, const int *data_length, const double *const *frequencies,
                           const double *const *data, const double *const *psd, const double *const *weights,
                           int *handle)
{
    if (num_detectors < 1 || !data_length || !frequencies || !psd || !handle)
        return gwh_set_error(GWH_ERR_INVALID, "null argument");
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    std::vector<const void *> ptrs;
    for (int d = 0; d < num_detectors; d++) {
        if (!frequencies[d] || !psd[d] || data_length[d] < 2) return gwh_set_error(GWH_ERR_INVALID, "bad arrays");
        ptrs.push_back(frequencies[d]);
        ptrs.push_back(psd[d]);
        if (data && data[d]) ptrs.push_back(data[d]);
        if (weights && weights[d]) ptrs.push_back(weights[d]);
    }
    *handle = E.next_reg++;
    E.host_regs.emplace_back(*handle, ptrs);
    bump_event_generation();
    return GWH_OK;
}

int gwh_host_data_release(int handle)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    for (size_t i = 0; i < E.host_regs.size(); i++)
        if (E.host_regs[i].first == handle) {
            E.host_regs.erase(E.host_regs.begin() + i);
            bump_event_generation();
            return E.ready ? drop_pointer_keyed_events(E) : GWH_OK;
        }
    return gwh_set_error(GWH_ERR_INVALID, "unknown host data handle");
}

int gwh_event_create_weighted(const double *frequencies, int length, int ndet, const int *detectors,
                              const double *data, const double *psd, int integration_method, const double *weights,
                              int log10F, int *event)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (integration_method != GWH_SIMPSONS && integration_method != GWH_GAUSSLEG)
        return gwh_set_error(GWH_ERR_INVALID, "integration method must be GWH_SIMPSONS or GWH_GAUSSLEG");
    return gwh_event_create_ex(frequencies, length, ndet, detectors, data, psd, integration_method,
                               integration_method == GWH_GAUSSLEG ? weights : nullptr, log10F ? 1 : 0, 0, 0, event);
}

int gwh_event_destroy(int event)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (!E.ready || event < 0 || event >= GWH_MAX_EVENTS || !E.events[event].live)
        return gwh_set_error(GWH_ERR_INVALID, "unknown event");
    if (E.events[event].pins > 0) return gwh_set_error(GWH_ERR_BUSY, "event is held by a step graph");
    GWH_HIP(hipStreamSynchronize(E.stream));
    GWH_HIP(hipFree(E.events[event].block));
    E.events[event] = Event();
    bump_event_generation();
    return GWH_OK;
}

size_t gwh_event_bytes(int event)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (!E.ready || event < 0 || event >= GWH_MAX_EVENTS || !E.events[event].live) return 0;
    return E.events[event].bytes;
}

int gwh_loglike_batch_device(const int *d_events, int events0, int model, int precision, const void *d_params, int n,
                             double *d_logl, void *stream)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    hipStream_t s = stream ? (hipStream_t)stream : E.stream;
    return gwh_loglike_launch(d_events, events0, model, precision, (const gwh_params *)d_params, n, d_logl, s, false,
                              nullptr);
}

int gwh_loglike_batch_device_noalloc(const int *d_events, int events0, int model, int precision, const void *d_params,
                                     int n, double *d_logl, void *stream)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    /* initialisation allocates the model tables: it must have happened before (gwh_init) */
    if (!E.ready || !E.noalloc_scratch)
        return gwh_set_error(GWH_ERR_INVALID, "gwh_loglike_batch_device_noalloc: call gwh_init first");
    if (n <= 0) return GWH_OK;
    /* the fixed array split: instance records, tile counts, then partial sums in the rest */
    const size_t a_coef = ((size_t)n * sizeof(InstCoef) + 255) & ~(size_t)255;
    const size_t a_act = ((size_t)n * sizeof(int) + 255) & ~(size_t)255;
    if (a_coef + a_act >= (size_t)GWH_NOALLOC_SCRATCH_BYTES)
        return gwh_set_error(GWH_ERR_NOMEM, std::to_string(n) + " instances exceed the allocation-free scratch of " +
                                                std::to_string(GWH_NOALLOC_SCRATCH_BYTES) + " bytes");
    Scratch S;
    unsigned char *base = E.noalloc_scratch;
    S.coef.p = base;
    S.coef.cap = a_coef;
    S.active.p = base + a_coef;
    S.active.cap = a_act;
    S.partial.p = base + a_coef + a_act;
    S.partial.cap = (size_t)GWH_NOALLOC_SCRATCH_BYTES - a_coef - a_act;
    for (DevBuf *b : {&S.coef, &S.active, &S.partial, &S.work, &S.wcount}) b->fixed = true;
    hipStream_t s = stream ? (hipStream_t)stream : E.stream;
    return gwh_loglike_launch(d_events, events0, model, precision, (const gwh_params *)d_params, n, d_logl, s, false,
                              nullptr, nullptr, 0, &S, nullptr);
}

int gwh_loglike_batch(const int *events, int events0, int model, int precision, const gwh_params *params, int n,
                      double *logl)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (n <= 0) return GWH_OK;
    if (!params || !logl) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    int rc;
    if ((rc = E.params.ensure((size_t)n * sizeof(gwh_params))) != GWH_OK) return rc;
    if ((rc = E.logl.ensure((size_t)n * sizeof(double))) != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(E.params.p, params, (size_t)n * sizeof(gwh_params), hipMemcpyHostToDevice, E.stream));
    const int *dev = nullptr;
    if (events) {
        for (int i = 0; i < n; i++)
            if (events[i] < 0 || events[i] >= GWH_MAX_EVENTS || !E.events[events[i]].live)
                return gwh_set_error(GWH_ERR_INVALID, "unknown event in batch");
        if ((rc = E.inst_event.ensure((size_t)n * sizeof(int))) != GWH_OK) return rc;
        GWH_HIP(hipMemcpyAsync(E.inst_event.p, events, (size_t)n * sizeof(int), hipMemcpyHostToDevice, E.stream));
        dev = E.inst_event.as<const int>();
        events0 = events[0];
    }
    rc = gwh_loglike_launch(dev, events0, model, precision, E.params.as<const gwh_params>(), n, E.logl.as<double>(),
                            E.stream, false, nullptr, events);
    if (rc != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(logl, E.logl.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

int gwh_time_loglike_batch_device(int events0, int model, int precision, const void *d_params, int n, double *d_logl,
                                  int reps, double *ms_per_launch, double *kernel_ms)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (reps < 1) reps = 1;
    /* one untimed launch: first-touch of the batch's buffers */
    int rc0 = gwh_loglike_launch(nullptr, events0, model, precision, (const gwh_params *)d_params, n, d_logl, E.stream,
                                 false, nullptr);
    if (rc0 != GWH_OK) return rc0;
    hipEvent_t a, b;
    GWH_HIP(hipEventCreate(&a));
    GWH_HIP(hipEventCreate(&b));
    double ksum = 0;
    GWH_HIP(hipEventRecord(a, E.stream));
    for (int r = 0; r < reps; r++) {
        double km = 0;
        int rc = gwh_loglike_launch(nullptr, events0, model, precision, (const gwh_params *)d_params, n, d_logl,
                                    E.stream, true, &km);
        if (rc != GWH_OK) return rc;
        ksum += km;
    }
    GWH_HIP(hipEventRecord(b, E.stream));
    GWH_HIP(hipEventSynchronize(b));
    float ms = 0;
    GWH_HIP(hipEventElapsedTime(&ms, a, b));
    GWH_HIP(hipEventDestroy(a));
    GWH_HIP(hipEventDestroy(b));
    if (ms_per_launch) *ms_per_launch = ms / reps;
    if (kernel_ms) *kernel_ms = ksum / reps;
    return GWH_OK;
}

int gwh_loglike_multi_device(const int *events, const int *d_events, int model, int precision, const void *d_params,
                             int n, double *d_logl, void *stream)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (n <= 0) return GWH_OK;
    if (!events || !d_events || !d_params || !d_logl) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    for (int i = 0; i < n; i++)
        if (events[i] < 0 || events[i] >= GWH_MAX_EVENTS || !E.events[events[i]].live)
            return gwh_set_error(GWH_ERR_INVALID, "unknown event in batch");
    hipStream_t s = stream ? (hipStream_t)stream : E.stream;
    return gwh_loglike_launch(d_events, events[0], model, precision, (const gwh_params *)d_params, n, d_logl, s, false,
                              nullptr, events);
}

/* Search over the tiling of each event-size bucket (floor(log2 N)): for every bucket present, time the whole batch
 * with that bucket's events tiled at each candidate (lanes per tile x bins per lane; the other buckets held), keep
 * the fastest (median of `reps` launches), and record it in the engine's tables for later events of the same size.
 * lanes_cand == nullptr: bins per lane only, each event keeping its lanes.  The tables (optional, 32 ints) receive
 * the choice per bucket (0 = absent). */
static int autotune_tiling(const int *events, const int *d_events, int model, int precision, const void *d_params,
                           int n, double *d_logl, int reps, const int *lanes_cand, int nlanes, int *bpl_table,
                           int *lanes_table)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (n <= 0 || !events || !d_events || !d_params || !d_logl) return gwh_set_error(GWH_ERR_INVALID, "bad arguments");
    if (reps < 1) reps = 1;
    std::vector<int> bucket_of(GWH_MAX_EVENTS, -1);
    bool present[32] = {false};
    for (int i = 0; i < n; i++) {
        int e = events[i];
        if (e < 0 || e >= GWH_MAX_EVENTS || !E.events[e].live)
            return gwh_set_error(GWH_ERR_INVALID, "unknown event in batch");
        if (E.events[e].pins > 0) return gwh_set_error(GWH_ERR_BUSY, "event is held by a step graph");
        int N = E.events[e].N, lg = 0;
        while ((2 << lg) <= N && lg < 31) lg++;
        bucket_of[e] = lg;
        present[lg] = true;
    }
    /* lanes <= 0: keep each event's lanes */
    auto retile = [&](int lg, int b, int lanes) -> int {
        for (int e = 0; e < GWH_MAX_EVENTS; e++) {
            if (bucket_of[e] != lg) continue;
            E.events[e].desc.bpl = b;
            if (lanes > 0) E.events[e].desc.lanes = lanes;
            GWH_HIP(hipMemcpyAsync(E.d_events + e, &E.events[e].desc, sizeof(EventDesc), hipMemcpyHostToDevice,
                                   E.stream));
        }
        return GWH_OK;
    };
    /* a batch on one event is timed as such a batch runs: the tile-major grid, not the work lists */
    bool one_event = true;
    for (int i = 1; i < n && one_event; i++) one_event = events[i] == events[0];
    auto time_batch = [&](double *ms) -> int {
        std::vector<float> t(reps);
        for (int r = 0; r < reps; r++) {
            GWH_HIP(hipEventRecord(E.ev0, E.stream));
            int rc = gwh_loglike_launch(one_event ? nullptr : d_events, events[0], model, precision,
                                        (const gwh_params *)d_params, n, d_logl, E.stream, false, nullptr,
                                        one_event ? nullptr : events);
            if (rc != GWH_OK) return rc;
            GWH_HIP(hipEventRecord(E.ev1, E.stream));
            GWH_HIP(hipEventSynchronize(E.ev1));
            GWH_HIP(hipEventElapsedTime(&t[r], E.ev0, E.ev1));
        }
        std::sort(t.begin(), t.end());
        *ms = t[reps / 2];
        return GWH_OK;
    };
    static const int cand[4] = {1, 2, 4, 8};
    static const int keep[1] = {0};
    if (!lanes_cand) {
        lanes_cand = keep;
        nlanes = 1;
    }
    int rc;
    double warm;
    if ((rc = time_batch(&warm)) != GWH_OK) return rc; /* first-touch and code-load warm-up */
    for (int lg = 0; lg < 32; lg++) {
        if (!present[lg]) continue;
        double best = 1e300;
        int bestb = 8, bestl = 0;
        for (int li = 0; li < nlanes; li++)
            for (int b : cand) {
                if ((rc = retile(lg, b, lanes_cand[li])) != GWH_OK) return rc;
                double ms;
                if ((rc = time_batch(&ms)) != GWH_OK) return rc;
                if (ms < best) {
                    best = ms;
                    bestb = b;
                    bestl = lanes_cand[li];
                }
            }
        E.tune[lg] = bestb;
        if (bestl > 0) E.tune_lanes[lg] = bestl;
        if ((rc = retile(lg, bestb, bestl)) != GWH_OK) return rc;
    }
    GWH_HIP(hipStreamSynchronize(E.stream));
    for (int lg = 0; lg < 32; lg++) {
        if (bpl_table) bpl_table[lg] = present[lg] ? E.tune[lg] : 0;
        if (lanes_table) lanes_table[lg] = present[lg] ? (E.tune_lanes[lg] > 0 ? E.tune_lanes[lg] : GWH_TPB) : 0;
    }
    return GWH_OK;
}

int gwh_autotune_bins_per_lane(const int *events, const int *d_events, int model, int precision, const void *d_params,
                               int n, double *d_logl, int reps, int *table)
{
    return autotune_tiling(events, d_events, model, precision, d_params, n, d_logl, reps, nullptr, 0, table, nullptr);
}

int gwh_autotune_tiles(const int *events, const int *d_events, int model, int precision, const void *d_params, int n,
                       double *d_logl, int reps, int *bpl_table, int *lanes_table)
{
    static const int lanes[5] = {64, 128, 256, 512, 1024};
    return autotune_tiling(events, d_events, model, precision, d_params, n, d_logl, reps, lanes, 5, bpl_table,
                           lanes_table);
}

int gwh_kernel_timing_start(int max_launches)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (max_launches < 1 || max_launches > (1 << 20)) return gwh_set_error(GWH_ERR_INVALID, "bad max_launches");
    while (E.kt_ev.size() < 2 * (size_t)max_launches) {
        hipEvent_t e;
        GWH_HIP(hipEventCreate(&e));
        E.kt_ev.push_back(e);
    }
    E.kt_used = 0;
    E.kt_on = true;
    E.kt_every = 1;
    E.kt_seq = 0;
    return GWH_OK;
}

int gwh_kernel_timing_every(int k)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (!E.kt_on) return gwh_set_error(GWH_ERR_INVALID, "kernel timing was not started");
    if (k < 1) return gwh_set_error(GWH_ERR_INVALID, "k must be >= 1");
    E.kt_every = k;
    E.kt_seq = 0;
    return GWH_OK;
}

int gwh_kernel_timing_stop(double *total_ms, int *launches)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (!E.kt_on) return gwh_set_error(GWH_ERR_INVALID, "kernel timing was not started");
    E.kt_on = false;
    double sum = 0;
    for (int i = 0; i < E.kt_used; i++) {
        GWH_HIP(hipEventSynchronize(E.kt_ev[2 * i + 1]));
        float ms = 0;
        GWH_HIP(hipEventElapsedTime(&ms, E.kt_ev[2 * i], E.kt_ev[2 * i + 1]));
        sum += ms;
    }
    if (total_ms) *total_ms = sum;
    if (launches) *launches = E.kt_used;
    return GWH_OK;
}

__global__ void gwh_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int gwh_device_copy(void *dst, const void *src, size_t bytes)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (!dst || !src || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return gwh_set_error(GWH_ERR_INVALID, "copy needs 16-byte aligned pointers and a multiple of 16 bytes");
    size_t n = bytes / 16;
    if (n == 0) return GWH_OK;
    hipLaunchKernelGGL(gwh_copy_kernel, dim3(4096), dim3(256), 0, E.stream, (const uint4 *)src, (uint4 *)dst, n);
    GWH_HIP(hipGetLastError());
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

/* ---- single-call drop-in entry points ---- */

static int upload_single(const double *f, int n, const gwh_params *p, int model, int det, const int *dets = nullptr,
                         int ndet = 1)
{
    Engine &E = gwh_engine();
    int rc;
    if ((rc = E.f.ensure((size_t)n * 8)) != GWH_OK) return rc;
    if ((rc = E.params.ensure(sizeof(gwh_params))) != GWH_OK) return rc;
    if ((rc = E.coef.ensure(sizeof(InstCoef))) != GWH_OK) return rc;
    if ((rc = E.out1.ensure((size_t)n * 16)) != GWH_OK) return rc;
    if ((rc = E.out2.ensure((size_t)n * 16)) != GWH_OK) return rc;
    if ((rc = E.misc.ensure(256)) != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(E.f.p, f, (size_t)n * 8, hipMemcpyHostToDevice, E.stream));
    GWH_HIP(hipMemcpyAsync(E.params.p, p, sizeof(gwh_params), hipMemcpyHostToDevice, E.stream));
    /* a one-detector descriptor for the setup (only the detector ids are read) */
    EventDesc ds;
    memset(&ds, 0, sizeof(ds));
    ds.D = ndet;
    ds.N = n;
    ds.bpl = 1;
    ds.lanes = GWH_TPB;
    for (int d = 0; d < GWH_MAXDET; d++) ds.det[d] = (dets && d < ndet) ? dets[d] : det;
    GWH_HIP(hipMemcpyAsync(E.d_events + (GWH_MAX_EVENTS - 1), &ds, sizeof(ds), hipMemcpyHostToDevice, E.stream));
    launch_setup(model, 1, E.stream, E.params.as<const gwh_params>(), nullptr, GWH_MAX_EVENTS - 1,
                 E.coef.as<InstCoef>(), nullptr);
    GWH_HIP(hipGetLastError());
    return GWH_OK;
}

static int check_common(const double *f, int n, const gwh_params *p, int model)
{
    if (!f || !p) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    if (n < 1) return gwh_set_error(GWH_ERR_INVALID, "length must be positive");
    if (!gwh_model_valid(model))
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "only IMRPhenomD, IMRPhenomPv2 and ppE IMRPhenomD are provided");
    return GWH_OK;
}

int gwh_fourier_waveform(const double *frequencies, int length, double *hplus, double *hcross, int model,
                         const gwh_params *params)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int rc = gwh_ensure_init();
    if (rc != GWH_OK) return rc;
    if ((rc = check_common(frequencies, length, params, model)) != GWH_OK) return rc;
    if (!hplus || !hcross) return gwh_set_error(GWH_ERR_INVALID, "null output");
    if ((rc = upload_single(frequencies, length, params, model, 0)) != GWH_OK) return rc;
    dim3 g((length + 255) / 256);
    if (!gwh_is_pv2(model))
        hipLaunchKernelGGL(gwh_waveform_kernel<GWH_IMRPHENOMD>, g, dim3(256), 0, E.stream, E.f.as<const double>(),
                           length, E.coef.as<const InstCoef>(), 0, E.out1.as<double2>(), E.out2.as<double2>());
    else
        hipLaunchKernelGGL(gwh_waveform_kernel<GWH_IMRPHENOMPV2>, g, dim3(256), 0, E.stream, E.f.as<const double>(),
                           length, E.coef.as<const InstCoef>(), 0, E.out1.as<double2>(), E.out2.as<double2>());
    GWH_HIP(hipGetLastError());
    GWH_HIP(hipMemcpyAsync(hplus, E.out1.p, (size_t)length * 16, hipMemcpyDeviceToHost, E.stream));
    GWH_HIP(hipMemcpyAsync(hcross, E.out2.p, (size_t)length * 16, hipMemcpyDeviceToHost, E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

int gwh_fourier_detector_response(const double *frequencies, int length, double *response, int detector, int model,
                                  const gwh_params *params)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int rc = gwh_ensure_init();
    if (rc != GWH_OK) return rc;
    if ((rc = check_common(frequencies, length, params, model)) != GWH_OK) return rc;
    if (!response) return gwh_set_error(GWH_ERR_INVALID, "null output");
    if (detector < 0 || detector >= GWH_NDET_TABLE)
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "unsupported detector (the reference exits here)");
    if ((rc = upload_single(frequencies, length, params, model, detector)) != GWH_OK) return rc;
    dim3 g((length + 255) / 256);
    if (!gwh_is_pv2(model))
        hipLaunchKernelGGL(gwh_waveform_kernel<GWH_IMRPHENOMD>, g, dim3(256), 0, E.stream, E.f.as<const double>(),
                           length, E.coef.as<const InstCoef>(), 1, E.out1.as<double2>(), E.out2.as<double2>());
    else
        hipLaunchKernelGGL(gwh_waveform_kernel<GWH_IMRPHENOMPV2>, g, dim3(256), 0, E.stream, E.f.as<const double>(),
                           length, E.coef.as<const InstCoef>(), 1, E.out1.as<double2>(), E.out2.as<double2>());
    GWH_HIP(hipGetLastError());
    GWH_HIP(hipMemcpyAsync(response, E.out1.p, (size_t)length * 16, hipMemcpyDeviceToHost, E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

/* the host arrays of a call, all registered (gwh_host_data_register)? */
static bool all_registered(const Engine &E, const std::vector<const void *> &ptrs)
{
    for (const void *p : ptrs) {
        bool found = false;
        for (auto &r : E.host_regs)
            for (const void *q : r.second) found = found || (q == p);
        if (!found) return false;
    }
    return true;
}

/* the arrays of one single call */
struct CallArrays {
    const double *const *f;
    int n, ndet;
    const int *dets;
    const double *const *data;
    const double *const *psd;
    int method;
    const double *weights;
    int log10F, df_mode;
    void pointers(std::vector<const void *> &ptrs) const
    {
        ptrs.clear();
        ptrs.push_back(f[0]);
        for (int d = 0; d < ndet; d++) {
            if (data) ptrs.push_back(data[d]);
            ptrs.push_back(psd[d]);
        }
        if (weights) ptrs.push_back(weights);
    }
    /* what the key depends on besides the arrays */
    uint64_t tail(uint64_t key) const
    {
        for (int d = 0; d < ndet; d++) key = gwh_hash_bytes(&dets[d], sizeof(int), key);
        key = gwh_hash_bytes(&method, sizeof(int), key);
        key = gwh_hash_bytes(&log10F, sizeof(int), key);
        return key == 0 ? 1 : key;
    }
    uint64_t pointer_key() const
    {
        std::vector<const void *> ptrs;
        pointers(ptrs);
        uint64_t key = 0x9a7e0000ull + (uint64_t)df_mode + (data ? 0 : 0x100);
        key = gwh_hash_bytes(ptrs.data(), ptrs.size() * sizeof(void *), key);
        key = gwh_hash_bytes(&n, sizeof(int), key);
        return tail(key);
    }
    uint64_t content_key() const
    {
        uint64_t key = 0x12345ull + (uint64_t)df_mode + (data ? 0 : 0x100);
        key = gwh_hash_bytes(f[0], (size_t)n * 8, key);
        for (int d = 0; d < ndet; d++) {
            if (data) key = gwh_hash_bytes(data[d], (size_t)n * 16, key);
            key = gwh_hash_bytes(psd[d], (size_t)n * 8, key);
        }
        if (weights) key = gwh_hash_bytes(weights, (size_t)n * 8, key);
        return tail(key);
    }
};

/* lk: the caller's hold on the engine lock, released while the arrays are hashed (content mode), so that
 * concurrent callers hash in parallel; null: the lock stays held */
static int single_call_event(const CallArrays &A, int *event, uint64_t *key_out, bool *by_ptr_out,
                             std::unique_lock<std::recursive_mutex> *lk)
{
    Engine &E = gwh_engine();
    std::vector<const void *> ptrs;
    A.pointers(ptrs);
    /* pointer key: the caller keeps these arrays unchanged while they are cached (pointer mode) or registered */
    const bool by_ptr = E.cache_by_pointer || (!E.host_regs.empty() && all_registered(E, ptrs));
    uint64_t key;
    if (by_ptr) {
        key = A.pointer_key();
    } else {
        if (lk) lk->unlock(); /* only the caller's arrays are read meanwhile */
        key = A.content_key();
        if (lk) lk->lock();
    }
    if (key_out) *key_out = key;
    if (by_ptr_out) *by_ptr_out = by_ptr;
    const int n = A.n, ndet = A.ndet;
    int live_cached = 0, oldest = -1;
    for (int i = 0; i < GWH_MAX_EVENTS - 1; i++) {
        Event &ev = E.events[i];
        if (!ev.live || ev.key == 0) continue;
        if (ev.key == key && ev.ptr_key == by_ptr && ev.N == n && ev.D == ndet) {
            ev.last_use = ++E.use_clock;
            *event = i;
            return GWH_OK;
        }
        live_cached++;
        if (ev.pins > 0) continue; /* held by a step graph or a call in flight: never evicted */
        if (oldest < 0 || ev.last_use < E.events[oldest].last_use) oldest = i;
    }
    if (live_cached >= 64 && oldest >= 0) { /* LRU of implicit events */
        GWH_HIP(hipStreamSynchronize(E.stream));
        GWH_HIP(hipFree(E.events[oldest].block));
        E.events[oldest] = Event();
        bump_event_generation();
    }
    std::vector<double> dat((size_t)ndet * n * 2, 0.0), ps((size_t)ndet * n);
    for (int d = 0; d < ndet; d++) {
        if (A.data) memcpy(&dat[(size_t)d * n * 2], A.data[d], (size_t)n * 16);
        memcpy(&ps[(size_t)d * n], A.psd[d], (size_t)n * 8);
    }
    int rc = gwh_event_create_ex(A.f[0], n, ndet, A.dets, dat.data(), ps.data(), A.method, A.weights, A.log10F,
                                 A.df_mode, key, event);
    if (rc == GWH_OK) E.events[*event].ptr_key = by_ptr;
    return rc;
}

int gwh_single_call_event(const double *const *f, int n, int ndet, const int *dets, const double *const *data,
                          const double *const *psd, int method, const double *weights, int log10F, int *event,
                          int df_mode)
{
    CallArrays A{f, n, ndet, dets, data, psd, method, weights, log10F, df_mode};
    return single_call_event(A, event, nullptr, nullptr, nullptr);
}

static int cached_event(const double *const *f, int n, int ndet, const int *dets, const double *const *data,
                        const double *const *psd, int method, const double *weights, int log10F, int *event,
                        int df_mode = 0)
{
    CallArrays A{f, n, ndet, dets, data, psd, method, weights, log10F, df_mode};
    return single_call_event(A, event, nullptr, nullptr, nullptr);
}

/* ---- single likelihood calls from many host threads ----
 * The reference's samplers call the likelihood from OpenMP and std::thread workers, one parameter set per call.
 * A call finds its resident event through a per-thread cache (no engine lock while the generation holds) and
 * queues its instance; a caller that finds a launch lane free leads: it takes every queued instance of the head's
 * model into one batch, launches it on its lane (engine lock held for the enqueue only), waits for its stream and
 * hands the results out; callers still queued meanwhile poll briefly, then block.  Two lanes let a second batch
 * start while the first is in flight.  An instance's result does not depend on its batch, so every call returns
 * what it would return alone. */
namespace {

struct EventLookup {
    uint64_t ptr_hash = 0, key = 0, gen = 0;
    int ev = -1;
    bool by_ptr = false;
};
constexpr int kLookups = 4;
thread_local EventLookup t_lookup[kLookups];
thread_local int t_lookup_next = 0;

std::atomic<bool> g_single_ready{false};

/* one queued call; its group is the likelihood model (one model per launch) */
struct SinglePending : gwh::CoalesceItem {
    int ev;
    uint64_t key;
    gwh_params p;
    double out = 0;
    int rc = GWH_OK;
};

constexpr int kSingleLanes = 2;
constexpr int kSingleMaxBatch = 4096;
constexpr int kStale = 1000; /* internal: the instance's event changed under a per-thread lookup; look up again */

/* A launch lane: its stream and scratch, and the batch's parameters, instance events and results in coherent pinned
 * host memory that the kernels read and write directly (no copies to enqueue for a batch of a few hundred bytes
 * per instance). */
struct SingleLane {
    hipStream_t stream = nullptr;
    Scratch scratch;
    gwh_params *h_params = nullptr, *g_params = nullptr; /* host / device view of the same memory */
    int *h_events = nullptr, *g_events = nullptr;
    double *h_logl = nullptr, *g_logl = nullptr;
    int cap = 0;
};

struct SingleCalls {
    gwh::Coalescer<kSingleLanes> co{kSingleMaxBatch};
    SingleLane lanes[kSingleLanes];
};
SingleCalls g_single;

int pinned_alloc(void **host, void **dev, size_t bytes)
{
    if (hipHostMalloc(host, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        *host = nullptr;
        return gwh_set_error(GWH_ERR_NOMEM, "pinned staging allocation failed");
    }
    if (hipHostGetDevicePointer(dev, *host, 0) != hipSuccess)
        return gwh_set_error(GWH_ERR_HIP, "no device address for pinned staging");
    return GWH_OK;
}

uint64_t pointer_tuple_hash(const CallArrays &A)
{
    std::vector<const void *> ptrs;
    A.pointers(ptrs);
    uint64_t h = gwh_hash_bytes(ptrs.data(), ptrs.size() * sizeof(void *), 0x51u + (uint64_t)A.df_mode);
    h = gwh_hash_bytes(&A.n, sizeof(int), h);
    h = gwh_hash_bytes(&A.ndet, sizeof(int), h);
    uint64_t flags = (A.data ? 1u : 0u);
    h = gwh_hash_bytes(&flags, sizeof(flags), h);
    return A.tail(h);
}

/* the resident event of a single call: per-thread cache at the current generation, else the engine's lookup */
int lookup_event(const CallArrays &A, int *ev, uint64_t *key)
{
    Engine &E = gwh_engine();
    const uint64_t ph = pointer_tuple_hash(A);
    const uint64_t gen = g_event_gen.load(std::memory_order_acquire);
    if (g_single_ready.load(std::memory_order_acquire)) {
        for (auto &L : t_lookup) {
            if (L.gen != gen || L.ptr_hash != ph || L.ev < 0) continue;
            if (L.by_ptr || A.content_key() == L.key) {
                *ev = L.ev;
                *key = L.key;
                return GWH_OK;
            }
        }
    }
    std::unique_lock<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    bool by_ptr = false;
    if ((st = single_call_event(A, ev, key, &by_ptr, &lk)) != GWH_OK) return st;
    EventLookup &L = t_lookup[t_lookup_next];
    t_lookup_next = (t_lookup_next + 1) % kLookups;
    L.ptr_hash = ph;
    L.key = *key;
    L.gen = g_event_gen.load(std::memory_order_acquire); /* read under the lock that every bump holds */
    L.ev = *ev;
    L.by_ptr = by_ptr;
    g_single_ready.store(true, std::memory_order_release);
    return GWH_OK;
}

void forget_lookups()
{
    for (auto &L : t_lookup) L = EventLookup();
}

/* enqueue and run one batch of queued calls on lane L (no coalescer lock held); `items` shrinks to the calls
 * launched (the others are marked stale) */
void run_single_batch(SingleLane &L, std::vector<gwh::CoalesceItem *> &items)
{
    Engine &E = gwh_engine();
    std::vector<SinglePending *> batch;
    for (auto *it : items) batch.push_back(static_cast<SinglePending *>(it));
    std::vector<int> evs;
    int n = 0, rc = GWH_OK;
    {
        std::lock_guard<std::recursive_mutex> lk(E.mu);
        rc = hipSetDevice(E.device) == hipSuccess ? GWH_OK : gwh_set_error(GWH_ERR_HIP, "hipSetDevice failed");
        if (rc == GWH_OK && !L.stream && hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess)
            rc = gwh_set_error(GWH_ERR_HIP, "hipStreamCreate failed");
        /* drop instances whose event changed since their lookup; the rest are pinned until the results are out */
        std::vector<SinglePending *> live;
        for (SinglePending *q : batch) {
            if (q->ev < 0 || q->ev >= GWH_MAX_EVENTS || !E.events[q->ev].live || E.events[q->ev].key != q->key) {
                q->rc = kStale;
                continue;
            }
            live.push_back(q);
        }
        batch.swap(live);
        items.assign(batch.begin(), batch.end());
        n = (int)batch.size();
        if (rc == GWH_OK && n > 0 && n > L.cap) {
            if (L.h_params) (void)hipHostFree(L.h_params);
            if (L.h_events) (void)hipHostFree(L.h_events);
            if (L.h_logl) (void)hipHostFree(L.h_logl);
            L.h_params = L.g_params = nullptr;
            L.h_events = L.g_events = nullptr;
            L.h_logl = L.g_logl = nullptr;
            L.cap = 0;
            const int cap = std::max(n, 256);
            if ((rc = pinned_alloc((void **)&L.h_params, (void **)&L.g_params, (size_t)cap * sizeof(gwh_params))) ==
                    GWH_OK &&
                (rc = pinned_alloc((void **)&L.h_events, (void **)&L.g_events, (size_t)cap * sizeof(int))) == GWH_OK &&
                (rc = pinned_alloc((void **)&L.h_logl, (void **)&L.g_logl, (size_t)cap * sizeof(double))) == GWH_OK)
                L.cap = cap;
        }
        bool multi = false;
        if (rc == GWH_OK && n > 0) {
            for (int i = 0; i < n; i++) {
                L.h_params[i] = batch[i]->p;
                L.h_events[i] = batch[i]->ev;
                multi = multi || (batch[i]->ev != batch[0]->ev);
                bool seen = false;
                for (int e : evs) seen = seen || (e == batch[i]->ev);
                if (!seen) evs.push_back(batch[i]->ev);
            }
            for (int e : evs) E.events[e].pins++;
            rc = gwh_loglike_launch(multi ? L.g_events : nullptr, batch[0]->ev, batch[0]->group, GWH_FP64,
                                    L.g_params, n, L.g_logl, L.stream, false, nullptr, multi ? L.h_events : nullptr, 0,
                                    &L.scratch);
        }
    }
    if (n > 0 && L.stream && hipStreamSynchronize(L.stream) != hipSuccess && rc == GWH_OK)
        rc = gwh_set_error(GWH_ERR_HIP, "stream synchronisation failed");
    for (int i = 0; i < n; i++) {
        batch[i]->rc = rc;
        batch[i]->out = (rc == GWH_OK) ? L.h_logl[i] : std::nan("");
    }
    if (!evs.empty()) {
        std::lock_guard<std::recursive_mutex> lk(E.mu);
        for (int e : evs)
            if (E.events[e].pins > 0) E.events[e].pins--;
    }
}

/* one FP64 instance through the coalescer; returns GWH_OK, an error, or kStale */
int coalesced_loglike(int ev, uint64_t key, int model, const gwh_params &p, double *out)
{
    SinglePending me;
    me.group = model;
    me.ev = ev;
    me.key = key;
    me.p = p;
    g_single.co.submit(me, [](int lane, std::vector<gwh::CoalesceItem *> &batch) {
        run_single_batch(g_single.lanes[lane], batch);
    });
    *out = me.out;
    return me.rc;
}

} // namespace

/* one FP64 instance of a single-call entry point: lookup, coalesced launch, one more lookup if the event changed */
static double single_call_loglike(const CallArrays &A, int model, const gwh_params &p)
{
    const double nan = std::nan("");
    for (int attempt = 0; attempt < 3; attempt++) {
        int ev;
        uint64_t key;
        if (lookup_event(A, &ev, &key) != GWH_OK) return nan;
        double out = nan;
        int rc = coalesced_loglike(ev, key, model, p, &out);
        if (rc == GWH_OK) return out;
        if (rc != kStale) return nan;
        forget_lookups();
    }
    gwh_set_error(GWH_ERR_BUSY, "the call's cached event kept changing");
    return nan;
}

int gwh_single_call_stats(double *stats)
{
    if (!stats) return gwh_set_error(GWH_ERR_INVALID, "null stats");
    long long launches, items, most;
    int queued;
    g_single.co.stats(&launches, &items, &most, &queued);
    for (int i = 0; i < 9; i++) stats[i] = 0;
    stats[0] = (double)launches;
    stats[1] = (double)items;
    stats[3] = (double)most;
    stats[6] = (double)queued;
    return GWH_OK;
}

double gwh_log_likelihood(const double *data, const double *psd, const double *frequencies, size_t length,
                          const gwh_params *params, int detector, int model)
{
    const double nan = std::nan("");
    int n = (int)length;
    if (check_common(frequencies, n, params, model) != GWH_OK || !data || !psd) {
        if (!data || !psd) gwh_set_error(GWH_ERR_INVALID, "null argument");
        return nan;
    }
    if (detector < 0 || detector >= GWH_NDET_TABLE) {
        gwh_set_error(GWH_ERR_UNSUPPORTED, "unsupported detector");
        return nan;
    }
    const double *fp[1] = {frequencies}, *dp[1] = {data}, *pp[1] = {psd};
    CallArrays A{fp, n, 1, &detector, dp, pp, GWH_SIMPSONS, nullptr, 0, 0};
    return single_call_loglike(A, model, *params);
}

double gwh_mcmc_likelihood_extrinsic(int save_waveform, const gwh_params *params, int model, const int *data_length,
                                     const double *const *frequencies, const double *const *data,
                                     const double *const *psd, const double *const *weights, int integration_method,
                                     int log10F, const int *detectors, int num_detectors, double T_eff)
{
    const double nan = std::nan("");
    if (!params || !data_length || !frequencies || !data || !psd || !detectors || num_detectors < 1) {
        gwh_set_error(GWH_ERR_INVALID, "null argument");
        return nan;
    }
    if (!gwh_model_valid(model)) {
        gwh_set_error(GWH_ERR_UNSUPPORTED, "unsupported model");
        return nan;
    }
    if (!save_waveform) {
        /* waveform_util.cpp:144-147: independent waveforms per detector are not supported */
        gwh_set_error(GWH_ERR_UNSUPPORTED, "Independent WFs for each detector currently not supported");
        return nan;
    }
    int n = data_length[0];
    if (n < 1) {
        gwh_set_error(GWH_ERR_INVALID, "length must be positive");
        return nan;
    }
    for (int d = 1; d < num_detectors; d++) {
        if (data_length[d] != n ||
            (frequencies[d] != frequencies[0] && memcmp(frequencies[d], frequencies[0], (size_t)n * 8) != 0)) {
            gwh_set_error(GWH_ERR_UNSUPPORTED, "all detectors must share one frequency grid");
            return nan;
        }
    }
    if (num_detectors > GWH_MAXDET) {
        gwh_set_error(GWH_ERR_UNSUPPORTED, "at most 4 detectors");
        return nan;
    }
    for (int d = 0; d < num_detectors; d++)
        if (detectors[d] < 0 || detectors[d] >= GWH_NDET_TABLE) {
            gwh_set_error(GWH_ERR_UNSUPPORTED, "unsupported detector");
            return nan;
        }
    const double *w = (integration_method == GWH_GAUSSLEG && weights) ? weights[0] : nullptr;
    CallArrays A{frequencies, n, num_detectors, detectors, data, psd, integration_method, w, log10F, 0};
    gwh_params p = *params;
    p.tc = T_eff - params->tc; /* mcmc_gw.cpp:2466-2468 */
    return single_call_loglike(A, model, p);
}

double gwh_log_likelihood_internal(const double *data, const double *psd, const double *frequencies,
                                   const double *weights, const double *response, int length, int log10F,
                                   int integration_method)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    const double nan = std::nan("");
    if (gwh_ensure_init() != GWH_OK) return nan;
    if (!data || !psd || !frequencies || !response || length < 2) {
        gwh_set_error(GWH_ERR_INVALID, "null argument or length < 2");
        return nan;
    }
    int det = 0, ev;
    const double *fp[1] = {frequencies}, *dp[1] = {data}, *pp[1] = {psd};
    if (cached_event(fp, length, 1, &det, dp, pp, integration_method,
                     integration_method == GWH_GAUSSLEG ? weights : nullptr, log10F, &ev) != GWH_OK)
        return nan;
    const EventDesc &ds = E.events[ev].desc;
    if (E.out1.ensure((size_t)length * 16) != GWH_OK) return nan;
    int nb = (length + GWH_TPB - 1) / GWH_TPB;
    if (nb > 1024) nb = 1024;
    if (E.partial.ensure((size_t)nb * 2 * 8) != GWH_OK) return nan;
    if (hipMemcpyAsync(E.out1.p, response, (size_t)length * 16, hipMemcpyHostToDevice, E.stream) != hipSuccess) {
        gwh_set_error(GWH_ERR_HIP, "copy failed");
        return nan;
    }
    hipLaunchKernelGGL(gwh_inner_kernel, dim3(nb), dim3(GWH_TPB), 0, E.stream, ds.dwS, E.out1.as<const double2>(),
                       ds.wS, length, E.partial.as<double>());
    std::vector<double> part(2 * nb);
    if (hipMemcpyAsync(part.data(), E.partial.p, part.size() * 8, hipMemcpyDeviceToHost, E.stream) != hipSuccess ||
        hipStreamSynchronize(E.stream) != hipSuccess) {
        gwh_set_error(GWH_ERR_HIP, "kernel failed");
        return nan;
    }
    double hh = 0, dh = 0;
    for (int b = 0; b < nb; b++) {
        hh += part[2 * b];
        dh += part[2 * b + 1];
    }
    return -0.5 * (hh - 2 * dh);
}

/* (a|b) sums of the one-detector event `ev` against a response: hh = (r|r), dh = (d|r) */
static int inner_products(int ev, const double *response, int length, double *hh, double *dh)
{
    Engine &E = gwh_engine();
    const EventDesc &ds = E.events[ev].desc;
    int rc;
    if ((rc = E.out1.ensure((size_t)length * 16)) != GWH_OK) return rc;
    int nb = (length + GWH_TPB - 1) / GWH_TPB;
    if (nb > 1024) nb = 1024;
    if ((rc = E.partial.ensure((size_t)nb * 2 * 8)) != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(E.out1.p, response, (size_t)length * 16, hipMemcpyHostToDevice, E.stream));
    hipLaunchKernelGGL(gwh_inner_kernel, dim3(nb), dim3(GWH_TPB), 0, E.stream, ds.dwS, E.out1.as<const double2>(),
                       ds.wS, length, E.partial.as<double>());
    GWH_HIP(hipGetLastError());
    std::vector<double> part(2 * nb);
    GWH_HIP(hipMemcpyAsync(part.data(), E.partial.p, part.size() * 8, hipMemcpyDeviceToHost, E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    double a = 0, b = 0;
    for (int k = 0; k < nb; k++) { /* fixed order over the block partials */
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    *hh = a;
    *dh = b;
    return GWH_OK;
}

double gwh_calculate_snr_internal(const double *psd, const double *waveform, const double *frequencies, int length,
                                  int integration_method, const double *weights, int log10F)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    const double nan = std::nan("");
    if (gwh_ensure_init() != GWH_OK) return nan;
    if (!psd || !waveform || !frequencies || length < 2 ||
        (integration_method == GWH_GAUSSLEG && !weights)) {
        gwh_set_error(GWH_ERR_INVALID, "null argument or length < 2");
        return nan;
    }
    std::vector<double> zeros((size_t)2 * length, 0.0);
    int det = 0, ev;
    const double *fp[1] = {frequencies}, *dp[1] = {zeros.data()}, *pp[1] = {psd};
    if (cached_event(fp, length, 1, &det, dp, pp, integration_method,
                     integration_method == GWH_GAUSSLEG ? weights : nullptr, log10F, &ev, 1) != GWH_OK)
        return nan;
    double hh, dh;
    if (inner_products(ev, waveform, length, &hh, &dh) != GWH_OK) return nan;
    return sqrt(hh);
}

double gwh_data_snr(const double *frequencies, int length, const double *data, const double *response,
                    const double *psd)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    const double nan = std::nan("");
    if (gwh_ensure_init() != GWH_OK) return nan;
    if (!frequencies || !data || !response || !psd || length < 2) {
        gwh_set_error(GWH_ERR_INVALID, "null argument or length < 2");
        return nan;
    }
    int det = 0, ev;
    const double *fp[1] = {frequencies}, *dp[1] = {data}, *pp[1] = {psd};
    if (cached_event(fp, length, 1, &det, dp, pp, GWH_SIMPSONS, nullptr, 0, &ev, 1) != GWH_OK) return nan;
    double hh, dh;
    if (inner_products(ev, response, length, &hh, &dh) != GWH_OK) return nan;
    return sqrt(dh);
}

int gwh_snr_batch(const int *events, int events0, int model, int precision, const gwh_params *params, int n,
                  double *snr)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int st = gwh_ensure_init();
    if (st != GWH_OK) return st;
    if (n <= 0) return GWH_OK;
    if (!params || !snr) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    int rc;
    if ((rc = E.params.ensure((size_t)n * sizeof(gwh_params))) != GWH_OK) return rc;
    if ((rc = E.logl.ensure((size_t)n * sizeof(double))) != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(E.params.p, params, (size_t)n * sizeof(gwh_params), hipMemcpyHostToDevice, E.stream));
    const int *dev = nullptr;
    if (events) {
        for (int i = 0; i < n; i++)
            if (events[i] < 0 || events[i] >= GWH_MAX_EVENTS || !E.events[events[i]].live)
                return gwh_set_error(GWH_ERR_INVALID, "unknown event in batch");
        if ((rc = E.inst_event.ensure((size_t)n * sizeof(int))) != GWH_OK) return rc;
        GWH_HIP(hipMemcpyAsync(E.inst_event.p, events, (size_t)n * sizeof(int), hipMemcpyHostToDevice, E.stream));
        dev = E.inst_event.as<const int>();
        events0 = events[0];
    }
    rc = gwh_loglike_launch(dev, events0, model, precision, E.params.as<const gwh_params>(), n, E.logl.as<double>(),
                            E.stream, false, nullptr, events, 1);
    if (rc != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(snr, E.logl.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

int gwh_work_counts(int event, int model, const gwh_params *params, int n, int *counts)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int rc = gwh_ensure_init();
    if (rc != GWH_OK) return rc;
    if (n <= 0 || !params || !counts) return gwh_set_error(GWH_ERR_INVALID, "bad arguments");
    if (event < 0 || event >= GWH_MAX_EVENTS || !E.events[event].live) return gwh_set_error(GWH_ERR_INVALID, "unknown event");
    if ((rc = E.params.ensure((size_t)n * sizeof(gwh_params))) != GWH_OK) return rc;
    if ((rc = E.coef.ensure((size_t)n * sizeof(InstCoef))) != GWH_OK) return rc;
    if ((rc = E.misc.ensure((size_t)n * GWH_NREGION_COUNTS * sizeof(int))) != GWH_OK) return rc;
    GWH_HIP(hipMemcpyAsync(E.params.p, params, (size_t)n * sizeof(gwh_params), hipMemcpyHostToDevice, E.stream));
    if ((rc = gwh_setup_launch(nullptr, event, model, E.params.as<const gwh_params>(), n, E.coef.as<InstCoef>(),
                               E.stream)) != GWH_OK)
        return rc;
    hipLaunchKernelGGL(gwh_region_count_kernel, dim3(n), dim3(GWH_TPB), 0, E.stream, E.coef.as<const InstCoef>(),
                       (const EventDesc *)E.d_events, event, E.misc.as<int>());
    GWH_HIP(hipGetLastError());
    GWH_HIP(hipMemcpyAsync(counts, E.misc.p, (size_t)n * GWH_NREGION_COUNTS * sizeof(int), hipMemcpyDeviceToHost,
                           E.stream));
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

int gwh_create_coherent_GW_detection(const int *detectors, int num_detectors, const double *frequencies, int length,
                                     int model, const gwh_params *params, double *responses)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int rc = gwh_ensure_init();
    if (rc != GWH_OK) return rc;
    if ((rc = check_common(frequencies, length, params, model)) != GWH_OK) return rc;
    if (!detectors || !responses || num_detectors < 1 || num_detectors > GWH_MAXDET)
        return gwh_set_error(GWH_ERR_INVALID, "1 to 4 detectors and an output buffer are required");
    for (int d = 0; d < num_detectors; d++)
        if (detectors[d] < 0 || detectors[d] >= GWH_NDET_TABLE)
            return gwh_set_error(GWH_ERR_UNSUPPORTED, "unsupported detector");
    if ((rc = upload_single(frequencies, length, params, model, detectors[0], detectors, num_detectors)) != GWH_OK)
        return rc;
    dim3 g((length + 255) / 256);
    for (int d = 0; d < num_detectors; d++) {
        if (!gwh_is_pv2(model))
            hipLaunchKernelGGL(gwh_waveform_kernel<GWH_IMRPHENOMD>, g, dim3(256), 0, E.stream, E.f.as<const double>(),
                               length, E.coef.as<const InstCoef>(), 2, E.out1.as<double2>(), E.out2.as<double2>(), d);
        else
            hipLaunchKernelGGL(gwh_waveform_kernel<GWH_IMRPHENOMPV2>, g, dim3(256), 0, E.stream,
                               E.f.as<const double>(), length, E.coef.as<const InstCoef>(), 2, E.out1.as<double2>(),
                               E.out2.as<double2>(), d);
        GWH_HIP(hipGetLastError());
        GWH_HIP(hipMemcpyAsync(responses + (size_t)2 * length * d, E.out1.p, (size_t)length * 16,
                               hipMemcpyDeviceToHost, E.stream));
    }
    GWH_HIP(hipStreamSynchronize(E.stream));
    return GWH_OK;
}

} /* extern "C" */
