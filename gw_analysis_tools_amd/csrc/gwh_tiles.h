/*
 * gwh_tiles.h -- the per-bin network log-likelihood kernels: one (instance, tile) work item per block, a tile of
 * TPB lanes x BPL bins per lane (TPB in 64..1024, BPL in 1..8, chosen per event: fixed default or autotuned), and
 * the mixed-precision pieces of the per-bin evaluation.  Included by the engine (declarations, the block reduction
 * and the work map) and by the four translation units that instantiate the kernels of one (model, precision) each,
 * so the 160 kernels compile in parallel.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "gwh_engine.h"

/* ------------------------------------------------------------------------- */
/* reductions                                                                 */
/* ------------------------------------------------------------------------- */

/* Block-reduce NV values per thread of a TPB-lane block (fixed order) and write them to out[0..nv); a one-wave
 * block writes its wave sums directly.  The block's LDS is not reused afterwards (no closing barrier). */
template <int NV, int TPB = GWH_TPB>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NV], double *__restrict__ out, double (*lds)[NV],
                                                   int nv = NV)
{
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; j++)
        if (j < nv) acc[j] = wave_sum(acc[j]); /* nv is uniform: unused slots are skipped, not masked */
    if (TPB == 64) {
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < NV; j++)
                if (j < nv) out[j] = acc[j];
        }
        return;
    }
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) lds[wv][j] = acc[j];
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        int j = threadIdx.x;
        double s = lds[0][j];
        for (int w = 1; w < TPB / 64; w++) s += lds[w][j];
        out[j] = s;
    }
}


/* ------------------------------------------------------------------------- */
/* per-bin mixed-precision evaluators                                         */
/* ------------------------------------------------------------------------- */

/* IMRPhenomD in mixed precision: the FP64 evaluator with the merger-ringdown amplitude's exponential in FP32
 * (TABLE_EXP false: ~6e-8 relative per bin) or through the FP64 table (true: a few ulp); the phase, including its
 * arctangent, stays FP64 (an FP32 atan there was the largest error of the path, ~7e-4 in logL for heavy
 * binaries). */
template <bool TABLE_EXP>
__device__ __forceinline__ void gwh_phenomd_mixed(const InstCoef &C, double f, const BinPow &bp, double &a, double &phi,
                                                  double &u_out, double &lx_out, const double2 *tab = gwh_turn_table,
                                                  const double *etab = gwh_exp2_table)
{
    double Mf = C.M * f;
    double r = C.Mm13 * bp.fm13;     /* (Mf)^{-1/3} */
    double u = C.M13 * bp.f13;       /* (Mf)^{1/3} */
    double r2 = r * r;
    double inv76 = C.Mm76 * bp.fm76; /* (Mf)^{-7/6} */
    double ar;
    if (f < C.f1) {
        double pn = C.ca[0] + u * u * (C.ca[2] + u * (C.ca[3] + u * (C.ca[4] + u * (C.ca[5] + u * C.ca[6]))));
        double u3 = u * u * u;
        ar = pn + u3 * u3 * u * (C.rho[0] + u * (C.rho[1] + u * C.rho[2]));
    } else if (f > C.f3) {
        double dl = f - C.fRD;
        ar = C.mr_k * (TABLE_EXP ? gwh_exp_table(-C.mr_g2 * dl, etab) : exp_fp32(-C.mr_g2 * dl)) * rcp_newton(dl * dl + C.mr_w2);
    } else {
        ar = C.delta[0] + Mf * (C.delta[1] + Mf * (C.delta[2] + Mf * (C.delta[3] + Mf * C.delta[4])));
    }
    a = C.amp_pref * ar * inv76;
    double psi;
    double lx = 0;
    if (f < C.f1p) {
        double v = u * 1.4645918875615231;
        lx = C.lgPiM + bp.lf; /* log(pi M f) */
        double c5 = C.c8 + lx * C.c9;
        double c6 = C.c10 - C.c11 * lx;
        double v2 = v * v;
        double poly = C.cp[0] + v2 * (C.cp[2] + v * (C.cp[3] + v * (C.cp[4] + v * (c5 + v * (c6 + v * C.cp[7])))));
        double iv = r * 0.68278406325529568; /* 1/v = pi^{-1/3} (Mf)^{-1/3} */
        double iv2 = iv * iv;
        double Mf43 = Mf * u;
        psi = -GWH_PI / 4. + C.K * poly * (iv2 * iv2 * iv) +
              (C.sig[0] * Mf + C.sig[1] * Mf43 + C.sig[2] * Mf43 * u + C.sig[3] * Mf * Mf);
        if (C.gimr) psi += gwh_gimr_bin(C, iv);
    } else if (f > C.f2p) {
        psi = C.alpha[0] + C.alpha[1] * Mf - C.alpha[2] * (r2 * r) + C.alpha[3] * (C.M34 * bp.f34) +
              C.alpha[4] * gwh_atan_seeded((f - C.alpha[5]) * C.inv_fdamp, tab);
    } else {
        double r3 = r2 * r;
        psi = C.beta[0] + C.beta[1] * Mf + C.beta[2] * (C.lgM + bp.lf) - C.beta[3] * (r3 * r3 * r3);
    }
    if (C.ppe_n > 0 && (C.ppe_imr || f < C.f1p)) psi += gwh_ppe_bin(C, bp.lf); /* FP64: phase accuracy */
    phi = psi - (C.tc_total * (f - C.f_ref) + C.phic);
    u_out = u;
    lx_out = lx;
}

/* ------------------------------------------------------------------------- */
/* the batched likelihood kernels                                             */
/* ------------------------------------------------------------------------- */

struct WorkMap {
    const int *inst_event; /* NULL: every instance uses event0 */
    int event0, n, tiles0; /* tiles0: tile stride of the partial buffer */
};

/* One (instance, tile) work item: TPB lanes x BPL bins of one instance's network log-likelihood, reduced
 * to the instance's partial sums for that tile.  Every early return is uniform over the block. */
template <int MODEL, int PREC, int BPL, int TPB>
__device__ __forceinline__ void gwh_loglike_tile(const WorkMap &W, const EventDesc *__restrict__ events,
                                                 const InstCoef *__restrict__ coefs, double *__restrict__ partial,
                                                 const int *__restrict__ active, const int inst, const int tile,
                                                 double (*lds)[GWH_PARTIAL_STRIDE], double2 *__restrict__ tab,
                                                 double *__restrict__ etab)
{
    {
        const int ev = W.inst_event ? W.inst_event[inst] : W.event0;
        const EventDesc &E = events[ev];
        if (E.bpl != BPL || E.lanes != TPB) return; /* the instance's event is tiled by another launch */
        if (tile >= active[inst]) return; /* above 0.2/M on an ascending grid: h = 0, nothing to add */
        const int w = inst * W.tiles0 + tile;
        const InstCoef &C = coefs[inst];
        const int N = E.N, D = E.D;
        double acc[GWH_PARTIAL_STRIDE];
#pragma unroll
        for (int j = 0; j < GWH_PARTIAL_STRIDE; j++) acc[j] = 0;
        /* skip tiles entirely above the 0.2/M cutoff (h = 0 there, IMRPhenomD.cpp:475) */
        constexpr int kLevel = gwh_tile_level(TPB * BPL);
        const double fmin_tile = gwh_g(E.tile_fmin)[E.fmin_off[kLevel] + tile];
        if (!(fmin_tile > C.fcut)) {
            /* the sin/cos and exponential tables in LDS: the detector phases (one gather per detector and bin), the
               merger-ringdown arctangent's correction and exponential and the IMRPhenomPv2 Euler angle read them
               there instead of through the vector memory path */
            for (int q = threadIdx.x; q < 256; q += TPB) tab[q] = gwh_turn_table[q];
            if (threadIdx.x < 64) etab[threadIdx.x] = gwh_exp2_table[threadIdx.x];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < BPL; j++) {
                /* Lanes past the end of the grid or above 0.2/M evaluate the last grid bin with a zero amplitude
                 * (and phase), so the accumulators never sit under divergent control flow; a wave without any
                 * bin at or below the cutoff skips the step on a uniform branch. */
                const int k0 = tile * TPB * BPL + j * TPB + threadIdx.x;
                const bool in = k0 < N;
                const int k = in ? k0 : N - 1;
                const double f = gwh_g(E.f)[k];
                const bool live = in && f <= C.fcut;
                if (!__ballot(live)) continue;
                const BinPow bp = gwh_load_bin(E, k);
                if (MODEL == GWH_IMRPHENOMD) {
                    double a, phi, u, lx;
                    /* mixed: the merger-ringdown exponential in FP32, everything else as FP64 */
                    gwh_phenomd_fp64<PREC == GWH_MIXED>(C, f, bp, a, phi, u, lx, tab, etab);
                    a = live ? a : 0.0;
                    phi = live ? phi : 0.0;
#pragma unroll
                    for (int d = 0; d < GWH_MAXDET; d++) {
                        if (d >= D) break;
                        double s, c;
                        gwh_sincos_tab(phi + C.det[d][3] * f, s, c, tab);
                        const double ws = gwh_g(E.wS)[(size_t)d * N + k];
                        const double2 dw = gwh_ld2(E.dwS, (size_t)d * N + k);
                        acc[3 * d] += a * a * ws;
                        acc[3 * d + 1] += a * (c * dw.x - s * dw.y);
                        acc[3 * d + 2] += a * (c * dw.y + s * dw.x);
                    }
                } else {
                    double A, Phi, hpr, hpi, hcr, hci;
                    if (PREC == GWH_FP64) {
                        gwh_pv2_fp64(C, f, bp, A, Phi, hpr, hpi, hcr, hci, tab, etab);
                    } else {
                        double u, lx;
                        gwh_phenomd_mixed<true>(C, f, bp, A, Phi, u, lx, tab, etab);
                        /* precession angles and the sin/cos of the Euler angle alpha in FP64 (alpha's rounding
                         * would enter |h|^2 directly), like the detector phases below: table-driven sin/cos and
                         * exponential (the FP32 ones left ~2.5e-4 in logL on short, loud precessing events) */
                        const double v = u * 1.4645918875615231;
                        const double logw = C.lgPiM + bp.lf; /* log(pi M f) */
                        const double x = v * v;
                        const double L = C.eta * (1.0 + C.L_c1 * x + C.L_c2 * x * x) * rcp_newton(v);
                        const double sv = C.SP * rcp_newton(L + C.SL);
                        const double cbd = rsqrt_seeded(1.0 + sv * sv);
                        /* sin(beta/2) = |s| cos b / (2 cos(beta/2)): no cancellation, one square root */
                        const double cbhd = sqrt(0.5 * (1.0 + cbd)), sbhd = fabs(sv) * cbd * (0.5 * rcp_newton(cbhd));
                        const double c2d = cbhd * cbhd, s2d = sbhd * sbhd;
                        const double d2[5] = {s2d * s2d, 2. * cbhd * s2d * sbhd, GWH_SQRT_6 * s2d * c2d,
                                              2. * c2d * cbhd * sbhd, c2d * c2d};
                        const double alpha = gwh_euler_eval(C.a_c, v, logw) + C.alpha_add;
                        const double eps = gwh_euler_eval(C.e_c, v, logw) + C.eps_add;
                        double sa, ca;
                        gwh_sincos_tab(alpha, sa, ca, tab);
                        double er[5], ei[5];
                        er[2] = 1.;
                        ei[2] = 0.;
                        er[3] = ca;
                        ei[3] = sa;
                        er[1] = ca;
                        ei[1] = -sa;
                        er[4] = ca * ca - sa * sa;
                        ei[4] = 2. * ca * sa;
                        er[0] = er[4];
                        ei[0] = -ei[4];
                        hpr = hpi = hcr = hci = 0;
#pragma unroll
                        for (int m = -2; m <= 2; m++) {
                            const double dm2 =
                                (m == -2) ? d2[4] : (m == -1) ? -d2[3] : (m == 0) ? d2[2] : (m == 1) ? -d2[1] : d2[0];
                            const double y = C.ylm[m + 2];
                            const double t2r = er[-m + 2] * dm2 * y, t2i = ei[-m + 2] * dm2 * y;
                            const double tmr = er[m + 2] * d2[m + 2] * y, tmi = ei[m + 2] * d2[m + 2] * y;
                            hpr += t2r + tmr;
                            hpi += t2i + tmi;
                            hcr -= t2i - tmi;
                            hci += t2r - tmr;
                        }
                        Phi = Phi + 2. * eps + C.t_corr_2pi * f;
                    }
                    A = live ? A : 0.0;
                    Phi = live ? Phi : 0.0;
#pragma unroll
                    for (int d = 0; d < GWH_MAXDET; d++) {
                        if (d >= D) break;
                        const double P = C.det[d][0], Q = C.det[d][1];
                        const double Gr = P * hpr + Q * hcr, Gi = P * hpi + Q * hci;
                        double s, c;
                        gwh_sincos_tab(Phi + C.det[d][3] * f, s, c, tab);
                        const double ws = gwh_g(E.wS)[(size_t)d * N + k];
                        const double2 dw = gwh_ld2(E.dwS, (size_t)d * N + k);
                        acc[3 * d] += A * A * (Gr * Gr + Gi * Gi) * ws;
                        acc[3 * d + 1] += A * (dw.x * (Gr * c + Gi * s) - dw.y * (Gr * s - Gi * c));
                    }
                }
            }
        }
        block_reduce_store<GWH_PARTIAL_STRIDE, TPB>(acc, partial + (size_t)w * GWH_PARTIAL_STRIDE, lds, 3 * D);
    }
}

/* the flattened (instance, tile) work list of one tiling class: one block per live work item, so batches mixing
 * 10^3- and 10^5-bin events launch no empty blocks; entries fold over y beyond 2^20 blocks */
template <int MODEL, int PREC, int BPL, int TPB>
__global__ __launch_bounds__(TPB) void gwh_loglike_wl_kernel(WorkMap W, const EventDesc *__restrict__ events,
                                                             const InstCoef *__restrict__ coefs,
                                                             double *__restrict__ partial,
                                                             const int *__restrict__ active,
                                                             const int2 *__restrict__ work,
                                                             const int *__restrict__ count)
{
    __shared__ double lds[TPB / 64][GWH_PARTIAL_STRIDE];
    __shared__ double2 tab[256];
    __shared__ double etab[64];
    const int w = blockIdx.y * gridDim.x + blockIdx.x;
    if (w >= *count) return;
    const int2 wi = work[w];
    gwh_loglike_tile<MODEL, PREC, BPL, TPB>(W, events, coefs, partial, active, wi.x, wi.y, lds, tab, etab);
}

/* One block per (instance, tile) item, in tile-major order (item w = tile * n + instance).  Consecutive blocks go
 * to different XCDs, so the instances of every tile spread over all XCDs; in instance-major order tile t of every
 * instance would land on the same XCD, and the tiles below an instance's 0.2/M cutoff -- the only ones with work --
 * on a few XCDs.  Items above the cutoff cost one load.  count (optional, device): only instances i < *count are
 * visited, so a launch sized for the largest possible count costs little when the count is small.
 * One item per block rather than a resident grid striding over the items: with a loop around the item the compiler
 * keeps loop-invariant values (constants of the polynomial chains, lane offsets) in registers across iterations,
 * which doubles the kernel's VGPRs (IMRPhenomD 66 -> 134, IMRPhenomPv2 90 -> 165) and halves the waves a SIMD
 * holds (7 -> 3 or 4); the results are bit-identical either way (512 instances on 16225 bins: FP64 IMRPhenomD
 * 36.4 -> 30.4 us, IMRPhenomPv2 60.6 -> 52.8 / mixed 58.2 -> 49.7 us). */
template <int MODEL, int PREC, int BPL, int TPB>
__global__ __launch_bounds__(TPB) void gwh_loglike_grid_kernel(WorkMap W, const EventDesc *__restrict__ events,
                                                               const InstCoef *__restrict__ coefs,
                                                               double *__restrict__ partial,
                                                               const int *__restrict__ active,
                                                               const int *__restrict__ count)
{
    __shared__ double lds[TPB / 64][GWH_PARTIAL_STRIDE];
    __shared__ double2 tab[256];
    __shared__ double etab[64];
    const int nn = count ? min(*count, W.n) : W.n;
    const int w = blockIdx.y * gridDim.x + blockIdx.x; /* items fold over y beyond 2^20 blocks */
    if (nn <= 0 || w >= nn * W.tiles0) return;
    gwh_loglike_tile<MODEL, PREC, BPL, TPB>(W, events, coefs, partial, active, w % nn, w / nn, lds, tab, etab);
}

/* ------------------------------------------------------------------------- */
/* launch of one tiling class                                                 */
/* ------------------------------------------------------------------------- */

template <int MODEL, int PREC, int BPL, int TPB>
static void gwh_launch_tiles_one(long long items, const int2 *work, hipStream_t st, const WorkMap &W,
                                 const EventDesc *ev, const InstCoef *c, double *part, const int *act,
                                 const int *count)
{
    /* a dispatch counts work-items in 32 bits per dimension: at most 2^20 blocks along x */
    const long long gx = std::max(1LL, std::min(items, 1LL << 20)), gy = std::max(1LL, (items + gx - 1) / gx);
    const dim3 g((unsigned)gx, (unsigned)gy);
    if (work)
        hipLaunchKernelGGL((gwh_loglike_wl_kernel<MODEL, PREC, BPL, TPB>), g, dim3(TPB), 0, st, W, ev, c, part, act,
                           work, count);
    else
        hipLaunchKernelGGL((gwh_loglike_grid_kernel<MODEL, PREC, BPL, TPB>), g, dim3(TPB), 0, st, W, ev, c, part,
                           act, count);
}

template <int MODEL, int PREC, int TPB>
static void gwh_launch_tiles_lanes(int bpl, long long items, const int2 *work, hipStream_t st, const WorkMap &W,
                                   const EventDesc *ev, const InstCoef *c, double *part, const int *act,
                                   const int *count)
{
    switch (bpl) {
    case 1: gwh_launch_tiles_one<MODEL, PREC, 1, TPB>(items, work, st, W, ev, c, part, act, count); break;
    case 2: gwh_launch_tiles_one<MODEL, PREC, 2, TPB>(items, work, st, W, ev, c, part, act, count); break;
    case 4: gwh_launch_tiles_one<MODEL, PREC, 4, TPB>(items, work, st, W, ev, c, part, act, count); break;
    default: gwh_launch_tiles_one<MODEL, PREC, 8, TPB>(items, work, st, W, ev, c, part, act, count); break;
    }
}

/* The per-bin kernel of one tiling class (bpl bins per lane x lanes per tile).  work == nullptr: one block per
 * (instance, tile) item of a batch over one event tiling (items = n x tiles; count = the optional device instance
 * count); otherwise one block per entry of the class's work list (items = its upper bound; count = its device
 * length).  Each (MODEL, PREC) is instantiated in its own translation unit (gwh_tiles_*.hip). */
template <int MODEL, int PREC>
void gwh_launch_tiles(int bpl, int lanes, long long items, const int2 *work, hipStream_t st, const WorkMap &W,
                      const EventDesc *ev, const InstCoef *c, double *part, const int *act, const int *count)
{
    switch (lanes) {
    case 64: gwh_launch_tiles_lanes<MODEL, PREC, 64>(bpl, items, work, st, W, ev, c, part, act, count); break;
    case 128: gwh_launch_tiles_lanes<MODEL, PREC, 128>(bpl, items, work, st, W, ev, c, part, act, count); break;
    case 512: gwh_launch_tiles_lanes<MODEL, PREC, 512>(bpl, items, work, st, W, ev, c, part, act, count); break;
    case 1024: gwh_launch_tiles_lanes<MODEL, PREC, 1024>(bpl, items, work, st, W, ev, c, part, act, count); break;
    default: gwh_launch_tiles_lanes<MODEL, PREC, 256>(bpl, items, work, st, W, ev, c, part, act, count); break;
    }
}
/* Load the code object holding the kernels of one (MODEL, PREC) on the current device: the runtime loads a
 * translation unit's kernels at the first use of any of them, a few milliseconds that would otherwise fall on the
 * first likelihood launch (and on a serving loop's first request). */
template <int MODEL, int PREC> int gwh_tiles_load()
{
    hipFuncAttributes a;
    return hipFuncGetAttributes(&a, reinterpret_cast<const void *>(&gwh_loglike_grid_kernel<MODEL, PREC, 8, 256>)) ==
                   hipSuccess
               ? 0
               : -1;
}
#define GWH_TILES_DECLARE(M, P)                                                                                    \
    extern template void gwh_launch_tiles<M, P>(int, int, long long, const int2 *, hipStream_t, const WorkMap &,    \
                                                const EventDesc *, const InstCoef *, double *, const int *,         \
                                                const int *);                                                       \
    extern template int gwh_tiles_load<M, P>();
GWH_TILES_DECLARE(GWH_IMRPHENOMD, GWH_FP64)
GWH_TILES_DECLARE(GWH_IMRPHENOMD, GWH_MIXED)
GWH_TILES_DECLARE(GWH_IMRPHENOMPV2, GWH_FP64)
GWH_TILES_DECLARE(GWH_IMRPHENOMPV2, GWH_MIXED)
