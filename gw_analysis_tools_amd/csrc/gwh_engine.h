/*
 * gwh_engine.h -- internal runtime of the engine (not part of the C ABI).
 *
 * One process-wide Engine per device holds the read-only model tables, the
 * resident events (frequency grid, Simpson/Gauss-Legendre weights folded with
 * 1/S, weighted data, tabulated powers of f) and grow-only scratch buffers, so the hot
 * path never calls hipMalloc once sizes have been seen.
 */
#ifndef GWH_ENGINE_H
#define GWH_ENGINE_H

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "gwh_model.h"

#define GWH_TPB 256          /* threads per block of the auxiliary kernels; the default lanes per likelihood tile */
#define GWH_NCLASS 20        /* tiling classes: bins per lane {1, 2, 4, 8} x lanes per tile {64, ..., 1024} */
#define GWH_FMIN_LEVELS 8    /* frequency-minimum levels: chunks of 64 * 2^L bins, L = 0..7 (64..8192 bins) */
#define GWH_MAX_EVENTS 4096
#define GWH_PARTIAL_STRIDE (3 * GWH_MAXDET)

/* Device-side view of one resident event. */
struct EventDesc {
    const double *f;        /* N */
    const double *wS;       /* D x N : integration weight * 4 / S  (FP64) */
    const double2 *dwS;     /* D x N : data * wS */
    const double *tile_fmin; /* smallest frequency of every chunk of 64 * 2^L bins, level L from fmin_off[L] */
    /* N each: log f, f^(1/3), f^(-1/3), f^(-7/6), f^(3/4) of every bin (gwh_bin_powers) */
    const double *lgf, *f13, *fm13, *fm76, *f34;
    int N, D, sorted; /* sorted: frequencies ascending */
    double snr_scale;         /* (f[1]-f[0]) / (weights' df): the SNR's Simpson step (waveform_util.cpp:497) */
    int bpl;                  /* bins per lane of this event's tiles (fixed per event: results never depend
                                 on which other events share a batch) */
    int lanes;                /* lanes per tile (64, 128, 256, 512 or 1024): a tile holds lanes x bpl bins */
    int fmin_off[GWH_FMIN_LEVELS];
    int det[GWH_MAXDET];
};

/* level of the frequency minima whose chunks are tiles of `bins` bins (a power of two in 64..8192) */
__host__ __device__ constexpr int gwh_tile_level(int bins)
{
    int l = 0;
    while ((64 << l) < bins) l++;
    return l;
}
/* (instance, tile) items of one instance on an event */
static inline int gwh_tiles_of(const EventDesc &d)
{
    const int t = d.lanes * d.bpl;
    return (d.N + t - 1) / t;
}

/* Loads through global-address-space pointers.  Pointers held in device structs are generic, and generic
 * (flat) loads count against the scalar-memory counter as well as the vector one, so every wait for a
 * scalar load of the instance coefficients would also wait for the bin data in flight. */
#define GWH_GLOBAL __attribute__((address_space(1)))
template <class T> __device__ __forceinline__ const GWH_GLOBAL T *gwh_g(const T *p) { return (const GWH_GLOBAL T *)p; }
__device__ __forceinline__ double2 gwh_ld2(const double2 *p, size_t k)
{
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 v = ((const GWH_GLOBAL v2 *)p)[k];
    return make_double2(v.x, v.y);
}

__device__ __forceinline__ BinPow gwh_load_bin(const EventDesc &E, int k)
{
    BinPow b;
    b.lf = gwh_g(E.lgf)[k];
    b.f13 = gwh_g(E.f13)[k];
    b.fm13 = gwh_g(E.fm13)[k];
    b.fm76 = gwh_g(E.fm76)[k];
    b.f34 = gwh_g(E.f34)[k];
    return b;
}

/* w_k of simpsons_sum (util.h:841-856): 1,4,2,4,...,(2|4),1 by index parity for any n */
__device__ __forceinline__ double simpson_w(int k, int n)
{
    if (k == 0 || k == n - 1) return (n == 1) ? 2.0 : 1.0;
    return (k & 1) ? 4.0 : 2.0;
}

/* lane move of an FP64 value by a DPP pattern (two 32-bit moves) */
template <int CTRL> __device__ __forceinline__ double dpp_move(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

/* Sum over the 64 lanes of a wave in a fixed order, with DPP lane moves on the vector ALU instead of
 * LDS permutes: pairs (quad_perm 1,0,3,2), quads (quad_perm 2,3,0,1), half rows (row_half_mirror), rows
 * (row_mirror); then the four row sums, read from lanes 0, 16, 32, 48, added as (r0 + r1) + (r2 + r3).
 * Every lane must be active; the result is uniform. */
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

struct Event {
    bool live = false;
    int N = 0, D = 0;
    int det[GWH_MAXDET] = {0, 0, 0, 0};
    void *block = nullptr; /* one allocation holding every array of the event */
    size_t bytes = 0;
    EventDesc desc;
    uint64_t key = 0;      /* key of the single-call cache (0 = explicit event): content or pointer key */
    bool ptr_key = false;  /* keyed by host pointers (pointer mode or registered arrays), not by content */
    double df_like = 0;    /* f[N/2] - f[N/2-1]: the likelihood's Simpson step (mcmc_gw.cpp:811) */
    double df_fisher = 0;  /* f[1] - f[0]: the Fisher's Simpson step (fisher.cpp:2746) */
    uint64_t last_use = 0;
    int pins = 0;          /* step graphs holding this event's buffers: no destroy, eviction or re-tiling */
    int method = 0, log10F = 0; /* integration rule of the stored weights */
};

/* grow-only device buffer; a fixed one is a view of memory the buffer does not own (never grown or freed:
 * a request beyond its capacity is an error) */
/* Capacity a growable device buffer of capacity `cap` holds after it is asked for `need` bytes: it keeps
 * what it has when that suffices, else it is reallocated with a quarter of headroom (4 KiB at least). */
static inline size_t gwh_grown_capacity(size_t need, size_t cap)
{
    if (need <= cap) return cap;
    return need < 4096 ? 4096 : need + need / 4;
}

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool fixed = false;
    int ensure(size_t bytes);
    void release();
    template <class T> T *as() { return (T *)p; }
};

/* Device scratch of one batched-likelihood launch sequence (setup, tiles, finalize).  The engine owns one;
 * the serving broker owns one per stream so batches can be in flight together. */
struct Scratch {
    DevBuf coef, partial, active, work, wcount;
    void release() { coef.release(); partial.release(); active.release(); work.release(); wcount.release(); }
};

/* Options of one batched-likelihood launch sequence beyond the common arguments.
 *   d_count       device int: only instances i < *d_count are evaluated (read when the kernels run)
 *   d_mask        device ints, n: only instances with d_mask[i] != 0 are evaluated
 *   reserve_only  size the scratch buffers for this launch and return without launching
 *   capturing     the stream is being captured into a graph: no timing events are recorded
 * Instances that are not evaluated get a NaN result. */
struct LaunchExtra {
    const int *d_count = nullptr;
    const int *d_mask = nullptr;
    bool reserve_only = false;
    bool capturing = false;
};

/* Device scratch and stream of one Fisher computation (the engine's own, or a broker worker's). */
struct FisherScratch {
    hipStream_t stream = nullptr;
    DevBuf params, coef, misc, deriv, gram, out, sig, raw;
    size_t max_operand = (size_t)1 << 30; /* operand rows of one pass: points per pass follow from it */
    /* pinned host staging of a pass's stencil records, weights and flags; two alternate, so the host builds the
     * next pass while the device runs the current one (stage_done: the copy out of each has run) */
    void *stage[2] = {nullptr, nullptr};
    size_t stage_cap[2] = {0, 0};
    hipEvent_t stage_done[2] = {nullptr, nullptr};
    size_t bytes() const /* device bytes */
    {
        return params.cap + coef.cap + misc.cap + deriv.cap + gram.cap + out.cap + sig.cap + raw.cap;
    }
    /* give back the device buffers (the stream's work finished first); all_host: the host staging too */
    void release(bool all_host = false)
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (DevBuf *d : {&params, &coef, &misc, &deriv, &gram, &out, &sig, &raw}) d->release();
        if (!all_host) return;
        for (int i = 0; i < 2; i++) {
            if (stage[i]) (void)hipHostFree(stage[i]);
            if (stage_done[i]) (void)hipEventDestroy(stage_done[i]);
            stage[i] = nullptr;
            stage_cap[i] = 0;
            stage_done[i] = nullptr;
        }
    }
};

struct Engine {
    std::recursive_mutex mu;
    bool ready = false;
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    DetGeom *d_geo = nullptr;
    double *d_qnm = nullptr;
    QnmTable qnm;
    EventDesc *d_events = nullptr;
    std::vector<Event> events;
    uint64_t use_clock = 0;
    int bpl = 0;   /* 0 = automatic */
    int lanes = 0; /* lanes per tile, 0 = automatic */
    int tune[32] = {0};       /* autotuned bins per lane by floor(log2 N) of the event grid (0 = default rule) */
    int tune_lanes[32] = {0}; /* autotuned lanes per tile by the same buckets (0 = default) */
    /* scratch */
    DevBuf params, coef, partial, active, logl, inst_event, f, data, psd, wts, out1, out2, misc;
    DevBuf work, wcount; /* flattened (instance, tile) lists of multi-event batches */
    FisherScratch fisher; /* stream = the engine stream */
    DevBuf fft, fftwork, mxpart, mxres, mxin; /* FFT-maximised likelihoods */
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    /* live timing of the per-bin kernel inside a caller's timed region: event pairs around each launch */
    std::vector<hipEvent_t> kt_ev;
    int kt_used = 0;
    bool kt_on = false;
    int kt_every = 1;        /* time every kt_every-th launch */
    long long kt_seq = 0;    /* launches since timing started */
    /* single calls: pointer-keyed cache mode, registered host arrays (handle -> arrays) */
    int cache_by_pointer = 0;
    std::vector<std::pair<int, std::vector<const void *>>> host_regs;
    int next_reg = 1;
    /* device address of the allocation-free entry point's scratch array on this engine's device
     * (hipGetSymbolAddress at initialisation; the array's name is not a device address in host code) */
    unsigned char *noalloc_scratch = nullptr;
};

Engine &gwh_engine();
int gwh_set_error(int code, const std::string &msg);
int gwh_ensure_init();

/* Create an event from host arrays. method: GWH_SIMPSONS / GWH_GAUSSLEG (weights used), log10F as in
 * Log_Likelihood_internal.  df_mode 0: the likelihood's conventions (Simpson step f[N/2]-f[N/2-1], the log10
 * Jacobian on Gauss-Legendre weights only); 1: the SNR's and Fisher's (f[1]-f[0], the Jacobian under either rule). */
int gwh_event_create_ex(const double *f, int N, int ndet, const int *dets, const double *data, const double *psd,
                        int method, const double *weights, int log10F, int df_mode, uint64_t key, int *event);
int gwh_choose_bpl(int N);
int gwh_choose_lanes(int N);

/* Resident event for the host arrays of a single-call entry point (engine lock held): a registered or pointer-mode
 * lookup by the arrays' addresses, else by a hash of their contents; created (and cached, LRU of 64) on a miss.
 * data == NULL: zero data. df_mode as for gwh_event_create_ex. */
extern "C" int gwh_single_call_event(const double *const *f, int n, int ndet, const int *dets,
                                     const double *const *data, const double *const *psd, int method,
                                     const double *weights, int log10F, int *event, int df_mode);

/* the batched likelihood on device buffers (engine lock held by caller) */
int gwh_loglike_launch(const int *d_events, int events0, int model, int precision, const gwh_params *d_params, int n,
                       double *d_logl, hipStream_t stream, bool timed, double *kernel_ms,
                       const int *h_events = nullptr, int what = 0, Scratch *scratch = nullptr,
                       const LaunchExtra *extra = nullptr);

/* Fisher matrices of n points on a resident event with the given scratch (no engine lock needed while
 * it runs; the event must stay alive).  The public entry points use the engine's scratch. */
int gwh_fisher_batch_with(int event, int model, int basis, int precision, const gwh_params *params, int n, int dim,
                          double epsilon, double *out, double *sigma, FisherScratch &F);

/* Device bytes scratch F holds after gwh_fisher_batch_with of these n points with at most max_operand bytes of
 * operand rows per pass (its buffers grow as gwh_grown_capacity says; 0 when the arguments are invalid, the
 * call itself reports why). */
size_t gwh_fisher_batch_bytes(int event, int model, int basis, int precision, const gwh_params *params, int n,
                              size_t max_operand, const FisherScratch &F, bool want_sigma);

/* per-instance setup on device buffers */
int gwh_setup_launch(const int *d_events, int events0, int model, const gwh_params *d_params, int n, InstCoef *d_coef,
                     hipStream_t stream);

uint64_t gwh_hash_bytes(const void *p, size_t n, uint64_t seed);

#define GWH_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return gwh_set_error(GWH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

#endif
