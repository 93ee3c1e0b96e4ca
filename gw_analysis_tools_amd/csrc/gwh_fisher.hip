/*
 * gwh_fisher.hip -- numerical Fisher matrices on MI355X.
 *
 *   host      unpack the point into the basis of the generation method, build the central-difference
 *             stencil (the perturbed parameter sets, fisher.cpp:148-557), run the FP64 setup for all of
 *             them in one launch
 *   stencil   one lane per bin and point: every perturbed waveform/response evaluated in FP64, the FP64
 *             difference quotients D_i(f) scaled by sqrt(w/S) written as rows of 16 (parameters padded
 *             to 16) -- the operand layout of the matrix-core Gram
 *   gram      F = sum_rows a a^T with v_mfma_f64_16x16x4 (FP64 path) or v_mfma_f32_16x16x4 (mixed path),
 *             64 rows of 4 per instruction per wave, FP64 sums across waves and row chunks
 *   invert    FP64 LU inverse per point (util.cpp:828-855) -> marginal sigmas
 *
 * Reference: src/fisher.cpp:81-145 (fisher_numerical), :148-557 (calculate_derivatives),
 * :1754-2572 (basis maps), :2704-2781 (calculate_fisher_elements).
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gwh_engine.h"
#include "gwh_tables.h"

typedef double gwh_d4 __attribute__((ext_vector_type(4)));
typedef float gwh_f4 __attribute__((ext_vector_type(4)));

#define GWH_FROWS 16      /* parameters padded to the MFMA tile */
#define GWH_MAXSTEN 4     /* order 4 uses 4 evaluations per parameter */
#define GWH_GRAM_ROWS 4096 /* operand rows per Gram block */
#define GWH_FUSED_BINS 64  /* bins per block of the fused stencil + Gram kernel: one per lane of each wave */
#define GWH_FUSED_MAX_POINTS 128 /* points per pass of the fused path (passes pipeline with the host staging) */
#ifndef GWH_FUSED_WAVES
#define GWH_FUSED_WAVES 8 /* waves per block of the fused kernel: GWH_FROWS / GWH_FUSED_WAVES operand rows each */
#endif

#include "gwh_basis.h"

static gwh_params non_parameter_options(const gwh_params &in)
{
    gwh_params w;
    memset(&w, 0, sizeof(w));
    w.chip = -1;
    w.phip = -1;
    w.sky_average = in.sky_average;
    w.f_ref = in.f_ref;
    w.gmst = in.gmst;
    w.shift_time = 0; /* fisher.cpp:2527 */
    w.shift_phase = in.shift_phase;
    w.dep_postmerger = in.dep_postmerger;
    /* fisher.cpp:2541-2570: the ppE terms and the gIMR lists travel with the parameters */
    w.ppe_nmod = in.ppe_nmod;
    w.ppe_theory = in.ppe_theory; /* a theory's map and redshift travel with its parameters (:2541-2545) */
    w.cosmology = in.cosmology;
    for (int i = 0; i < GWH_PPE_MAX; i++) {
        w.ppe_b[i] = in.ppe_b[i];
        w.ppe_beta[i] = in.ppe_beta[i];
    }
    memcpy(w.gimr_dphi, in.gimr_dphi, sizeof(w.gimr_dphi));
    memcpy(w.gimr_dsigma, in.gimr_dsigma, sizeof(w.gimr_dsigma));
    memcpy(w.gimr_dbeta, in.gimr_dbeta, sizeof(w.gimr_dbeta));
    memcpy(w.gimr_dalpha, in.gimr_dalpha, sizeof(w.gimr_dalpha));
    w.gimr_nlist = in.gimr_nlist;
    memcpy(w.gimr_list, in.gimr_list, sizeof(w.gimr_list));
    return w;
}

/* the table entry of a gIMR Fisher-basis code (GWH_GIMR_CODE_*), null for a parameter the waveform does not read */
static const double *gimr_slot(const gwh_params &p, int code)
{
    if (code < 0 || code >= GWH_GIMR_NCODES) return nullptr;
    if (code < GWH_GIMR_CODE_SIGMA(2)) return &p.gimr_dphi[code];
    if (code < GWH_GIMR_CODE_BETA(2)) return &p.gimr_dsigma[code - GWH_GIMR_CODE_SIGMA(2)];
    if (code < GWH_GIMR_CODE_ALPHA(2)) return &p.gimr_dbeta[code - GWH_GIMR_CODE_BETA(2)];
    return &p.gimr_dalpha[code - GWH_GIMR_CODE_ALPHA(2)];
}

/* Fisher basis of a model: the GR basis, with betappe[0..nmod) appended for the ppE methods
 * (fisher.cpp:2083-2088 unpack, :2424-2429 repack) or the listed gIMR deltas for the gIMR methods
 * (:2128-2158, :2473-2505; a repeated entry repacks into the table only through its first occurrence, the one
 * the waveform reads) */
static void unpack_mod(const gwh_params &p, int basis, int var, int nmod, double *v, int *lf)
{
    unpack(p, basis, v, lf);
    const int base = basis_dim(basis);
    for (int i = 0; i < nmod; i++) {
        const double *s = (var == 3) ? gimr_slot(p, p.gimr_list[i]) : &p.ppe_beta[i];
        v[base + i] = s ? *s : 0.0;
        lf[base + i] = 0;
    }
}
static void repack_mod(const double *v, int basis, int var, int nmod, gwh_params &q)
{
    repack(v, basis, q);
    const int base = basis_dim(basis);
    for (int i = 0; i < nmod; i++) {
        if (var != 3) {
            q.ppe_beta[i] = v[base + i];
            continue;
        }
        bool first = true;
        for (int j = 0; j < i; j++) first &= (q.gimr_list[j] != q.gimr_list[i]);
        double *s = const_cast<double *>(gimr_slot(q, q.gimr_list[i]));
        if (first && s) *s = v[base + i];
    }
}

/* appended parameters of a model's Fisher basis: the ppE betas or the listed gIMR deltas (every point of a
 * batch with the same count) */
static int fisher_mods(int model, const gwh_params *params, int n, int *nmod)
{
    *nmod = 0;
    if (!gwh_model_valid(model)) return gwh_set_error(GWH_ERR_UNSUPPORTED, "unknown model");
    const int var = gwh_ppe_variant(model);
    if (var == 0 || n <= 0) return GWH_OK;
    const int m = (var == 3) ? params[0].gimr_nlist : params[0].ppe_nmod;
    if (var != 3 && (m < 0 || m > GWH_PPE_MAX))
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "ppE term count outside 0..4");
    if (var == 3 && (m < 0 || m > GWH_GIMR_LIST_MAX))
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "at most 16 gIMR deltas in a Fisher basis");
    for (int i = 0; i < n; i++) {
        if (((var == 3) ? params[i].gimr_nlist : params[i].ppe_nmod) != m)
            return gwh_set_error(GWH_ERR_INVALID, "every point of a modified-gravity Fisher batch needs the same "
                                                  "number of ppE terms / gIMR deltas");
        if (var == 3)
            for (int k = 0; k < m; k++)
                if (params[i].gimr_list[k] < -1 || params[i].gimr_list[k] >= GWH_GIMR_NCODES)
                    return gwh_set_error(GWH_ERR_INVALID, "bad gIMR Fisher-basis code");
    }
    *nmod = m;
    return GWH_OK;
}

static double host_dtoa(double ra, double dec, double gmst, int d1, int d2)
{
    const double *l1 = gwh_det_location[d1], *l2 = gwh_det_location[d2];
    double ha = gmst - ra;
    double e0 = cos(dec) * cos(ha), e1 = cos(dec) * -sin(ha), e2 = sin(dec);
    return ((l1[0] - l2[0]) * e0 + (l1[1] - l2[1]) * e1 + (l1[2] - l2[2]) * e2) / GWH_C_SI;
}

/* Stencil of one point: for every parameter i, S evaluations with weights c[i][s] (the difference
 * quotient including 1/eps and the log factor theta_i); instance parameter sets are appended to
 * `inst`, `no_tof` marks the order-4 +-2 eps evaluations that shift tc instead of using the time-of-flight
 * phase (fisher.cpp:433-453). */
struct PointStencil {
    int S;
    double c[GWH_FROWS][GWH_MAXSTEN];
};

/* Step of one basis coordinate v: step > 0 is the absolute epsilon of fisher.cpp:158; step < 0 is a
 * step relative to the coordinate, |step| * |v| (|step| where v is exactly 0). */
static double coordinate_step(double step, double v)
{
    if (step > 0) return step;
    return v != 0 ? -step * fabs(v) : -step;
}

static int build_stencil(const gwh_params &p, int basis, int var, int nmod, int order, double step, int det,
                         int ref_det, std::vector<gwh_params> &inst, std::vector<unsigned char> &no_tof,
                         PointStencil &ps)
{
    int dim = basis_dim(basis) + nmod;
    double v[GWH_FROWS];
    int lf[GWH_FROWS];
    unpack_mod(p, basis, var, nmod, v, lf);
    gwh_params base = non_parameter_options(p);
    ps.S = (order >= 4) ? 4 : 2;
    for (int i = 0; i < GWH_FROWS; i++)
        for (int s = 0; s < GWH_MAXSTEN; s++) ps.c[i][s] = 0;
    for (int i = 0; i < dim; i++) {
        double eps = coordinate_step(step, v[i]);
        bool edge = (i == 8 && v[i] > .25 - eps); /* eta at the equal-mass boundary: one-sided */
        double fac = lf[i] ? v[i] : 1.0;
        double off[4];
        int S = ps.S;
        if (S == 2) {
            off[0] = edge ? 0 : eps;
            off[1] = -eps;
            double den = edge ? eps : 2 * eps;
            ps.c[i][0] = fac / den;
            ps.c[i][1] = -fac / den;
        } else {
            off[0] = edge ? 0 : 2 * eps;
            off[1] = eps;
            off[2] = -eps;
            off[3] = -2 * eps;
            double den = edge ? 6. * eps : 12. * eps;
            ps.c[i][0] = -fac / den;
            ps.c[i][1] = 8. * fac / den;
            ps.c[i][2] = -8. * fac / den;
            ps.c[i][3] = fac / den;
        }
        for (int s = 0; s < S; s++) {
            double w[GWH_FROWS];
            for (int j = 0; j < dim; j++) w[j] = v[j];
            w[i] = v[i] + off[s];
            gwh_params q = base;
            repack_mod(w, basis, var, nmod, q);
            bool two = (S == 4) && (s == 0 || s == 3);
            if (two && det != ref_det) q.tc -= host_dtoa(q.RA, q.DEC, q.gmst, ref_det, det);
            inst.push_back(q);
            no_tof.push_back(two ? 1 : 0);
        }
    }
    return dim;
}

/* ------------------------------------------------------------------------- */
/* device: stencil                                                            */
/* ------------------------------------------------------------------------- */

struct StencilArgs {
    const InstCoef *coef;        /* points x (dim*S [+1 central]) records */
    const unsigned char *no_tof;
    const double *cw;            /* points x 16 x 4 stencil weights */
    int dim, S, per_point;       /* records per point */
    int d_first, d_count;        /* detectors of the event accumulated */
    int skyavg;                  /* sky-averaged IMRPhenomD: amplitude / phase derivatives */
    double sa_sign;              /* +1 (order 2) or -1 (order 4): i d(phase) A sign of fisher.cpp:299-315 */
    double *raw;                 /* non-null: write the bare derivatives dim x N complex (one point, one
                                    detector) instead of weighted Gram operand rows */
};

/* One lane per (bin, point, operand row): blockIdx.z is the row, and the S = 2 or 4 stencil evaluations of that
 * component are unrolled (S = a.S, a template parameter).  With the rows and stencil points as loops inside the
 * lane the compiler kept loop-invariant values live across them (IMRPhenomD 154, IMRPhenomPv2 190 VGPRs: 3 and
 * 2 waves per SIMD); now 115-127 VGPRs, 4 waves.  The arithmetic is unchanged. */
template <int MODEL, typename T, int S>
__global__ __launch_bounds__(GWH_TPB) void gwh_fisher_stencil_kernel(StencilArgs a, const EventDesc *__restrict__ events,
                                                                     int event0, T *__restrict__ rows)
{
    const int pt = blockIdx.y;
    const int i = blockIdx.z; /* operand row: component i < dim, zero padding up to GWH_FROWS */
    const int k = blockIdx.x * GWH_TPB + threadIdx.x;
    const EventDesc &E = events[event0];
    const int N = E.N;
    if (k >= N) return;
    const double f = E.f[k];
    const InstCoef *C = a.coef + (size_t)pt * a.per_point;
    const double *cw = a.cw + (size_t)pt * GWH_FROWS * GWH_MAXSTEN;
    const unsigned char *nt = a.no_tof + (size_t)pt * a.per_point;
    const size_t nrows = (size_t)a.d_count * N * 2;
    T *base = rows + (size_t)pt * nrows * GWH_FROWS;
    if (a.skyavg) {
        /* d h = dA + i sign d(phase) A_central (fisher.cpp:292-316); no detector projection */
        const InstCoef &Cc = C[a.dim * S];
        const BinPow bp = gwh_load_bin(E, k);
        double Ac = 0, u, lx, ph;
        if (f <= Cc.fcut) gwh_phenomd_fp64(Cc, f, bp, Ac, ph, u, lx);
        double sw = sqrt(E.wS[(size_t)a.d_first * N + k]);
        double dA = 0, dP = 0;
        if (i < a.dim) {
#pragma unroll
            for (int s = 0; s < S; s++) {
                const InstCoef &Ci = C[i * S + s];
                double A = 0, phi;
                gwh_phenomd_fp64(Ci, f, bp, A, phi, u, lx);
                if (!(f <= Ci.fcut)) A = 0;
                dA += cw[i * GWH_MAXSTEN + s] * A;
                dP += cw[i * GWH_MAXSTEN + s] * (-phi); /* construct_phase returns -(Psi - tc (f-fref) - phic) */
            }
        }
        if (a.raw) {
            if (i < a.dim) {
                a.raw[((size_t)i * N + k) * 2] = dA;
                a.raw[((size_t)i * N + k) * 2 + 1] = a.sa_sign * dP * Ac;
            }
            return;
        }
        base[(size_t)i * N + k] = (T)(dA * sw);
        base[(size_t)(GWH_FROWS + i) * N + k] = (T)(a.sa_sign * dP * Ac * sw);
        return;
    }
    const BinPow bp = gwh_load_bin(E, k);
    double sw[GWH_MAXDET];
    for (int dd = 0; dd < GWH_MAXDET; dd++) sw[dd] = (dd < a.d_count) ? sqrt(E.wS[(size_t)(a.d_first + dd) * N + k]) : 0;
    double dr[GWH_MAXDET] = {0, 0, 0, 0}, di[GWH_MAXDET] = {0, 0, 0, 0};
    if (i < a.dim) {
#pragma unroll
        for (int s = 0; s < S; s++) {
            const InstCoef &Ci = C[i * S + s];
            if (!(f <= Ci.fcut)) continue; /* h = 0 above 0.2/M of this evaluation */
            const double c = cw[i * GWH_MAXSTEN + s];
            const bool tof = !nt[i * S + s];
            if (MODEL == GWH_IMRPHENOMD) {
                double A, phi, u, lx;
                gwh_phenomd_fp64(Ci, f, bp, A, phi, u, lx);
                for (int dd = 0; dd < GWH_MAXDET; dd++) {
                    if (dd >= a.d_count) break;
                    int d = a.d_first + dd;
                    double sn, cs;
                    gwh_sincos_table(phi + (tof ? Ci.det[d][3] : 0.0) * f, sn, cs);
                    double er = A * cs, ei = -A * sn; /* A K e^{-i(phi + 2 pi dt f)} */
                    dr[dd] += c * (er * Ci.det[d][0] - ei * Ci.det[d][1]);
                    di[dd] += c * (er * Ci.det[d][1] + ei * Ci.det[d][0]);
                }
            } else {
                double A, Phi, hpr, hpi, hcr, hci;
                gwh_pv2_fp64(Ci, f, bp, A, Phi, hpr, hpi, hcr, hci);
                for (int dd = 0; dd < GWH_MAXDET; dd++) {
                    if (dd >= a.d_count) break;
                    int d = a.d_first + dd;
                    double Gr = Ci.det[d][0] * hpr + Ci.det[d][1] * hcr;
                    double Gi = Ci.det[d][0] * hpi + Ci.det[d][1] * hci;
                    double sn, cs;
                    gwh_sincos_table(Phi + (tof ? Ci.det[d][3] : 0.0) * f, sn, cs);
                    dr[dd] += c * A * (Gr * cs + Gi * sn);
                    di[dd] += c * A * (Gi * cs - Gr * sn);
                }
            }
        }
    }
    if (a.raw) {
        if (i < a.dim) {
            a.raw[((size_t)i * N + k) * 2] = dr[0];
            a.raw[((size_t)i * N + k) * 2 + 1] = di[0];
        }
        return;
    }
    for (int dd = 0; dd < GWH_MAXDET; dd++) {
        if (dd >= a.d_count) break;
        T *pl = base + (size_t)dd * 2 * GWH_FROWS * N; /* planes (detector, re/im), component-major */
        pl[(size_t)i * N + k] = (T)(dr[dd] * sw[dd]);
        pl[(size_t)(GWH_FROWS + i) * N + k] = (T)(di[dd] * sw[dd]);
    }
}

/* the stencil kernel of a model, operand type and stencil size (a.S: 2 for order 2, 4 for order 4) */
template <typename T>
static void launch_stencil(bool pv2, dim3 g, hipStream_t st, const StencilArgs &a, const EventDesc *evp, int event,
                           T *rows)
{
    if (!pv2) {
        if (a.S == 4)
            hipLaunchKernelGGL((gwh_fisher_stencil_kernel<GWH_IMRPHENOMD, T, 4>), g, dim3(GWH_TPB), 0, st, a, evp,
                               event, rows);
        else
            hipLaunchKernelGGL((gwh_fisher_stencil_kernel<GWH_IMRPHENOMD, T, 2>), g, dim3(GWH_TPB), 0, st, a, evp,
                               event, rows);
    } else {
        if (a.S == 4)
            hipLaunchKernelGGL((gwh_fisher_stencil_kernel<GWH_IMRPHENOMPV2, T, 4>), g, dim3(GWH_TPB), 0, st, a, evp,
                               event, rows);
        else
            hipLaunchKernelGGL((gwh_fisher_stencil_kernel<GWH_IMRPHENOMPV2, T, 2>), g, dim3(GWH_TPB), 0, st, a, evp,
                               event, rows);
    }
}

/* ------------------------------------------------------------------------- */
/* device: Gram on the matrix cores                                           */
/* ------------------------------------------------------------------------- */

/* The operand rows of a point lie in planes (detector, real / imaginary part), component-major: component i of
 * bin k at plane[i * N + k], so the stencil kernels write whole waves of consecutive bins.  Each block: up to
 * GWH_GRAM_ROWS bins of one plane of one point, 4 waves; each wave accumulates a 16x16 tile of a a^T over its
 * bins, 4 bins per MFMA (lane l supplies component l%16 of bin k0 + l/16 as both A and B operand; eight
 * MFMAs' operands are loaded first, 32 consecutive bins of each component).  Output: per block 256 FP64
 * partial sums. */
template <typename T>
__global__ __launch_bounds__(GWH_TPB) void gwh_gram_kernel(const T *__restrict__ rows, int N, int nplanes,
                                                           int blk_per_plane, double *__restrict__ partial, int nblk)
{
    __shared__ double lds[GWH_TPB / 64][256];
    const int pt = blockIdx.y, blk = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int plane = blk / blk_per_plane;
    const T *a = rows + ((size_t)pt * nplanes + plane) * GWH_FROWS * N + (size_t)(lane & 15) * N;
    int r0 = (blk % blk_per_plane) * GWH_GRAM_ROWS + wv * (GWH_GRAM_ROWS / 4);
    int r1 = r0 + GWH_GRAM_ROWS / 4;
    if (r1 > N) r1 = N;
    double out[4];
    if (sizeof(T) == 8) {
        gwh_d4 acc = {0, 0, 0, 0};
        for (int r = r0; r < r1; r += 32) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int rr = r + 4 * u + (lane >> 4);
                x[u] = (rr < r1) ? (double)a[rr] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x[u], x[u], acc, 0, 0, 0);
        }
        for (int j = 0; j < 4; j++) out[j] = acc[j];
    } else {
        gwh_f4 acc = {0, 0, 0, 0};
        for (int r = r0; r < r1; r += 32) {
            float x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int rr = r + 4 * u + (lane >> 4);
                x[u] = (rr < r1) ? (float)a[rr] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x[u], x[u], acc, 0, 0, 0);
        }
        for (int j = 0; j < 4; j++) out[j] = acc[j];
    }
    /* result layout on gfx950 (checked on the device): column l % 16; row l / 16 + 4 j for the FP64
     * instruction, 4 (l / 16) + j for the FP32 one */
    for (int j = 0; j < 4; j++) {
        int row = (sizeof(T) == 8) ? ((lane >> 4) + 4 * j) : (4 * (lane >> 4) + j);
        lds[wv][row * 16 + (lane & 15)] = out[j];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 256; e += GWH_TPB) {
        double s = lds[0][e];
        for (int w = 1; w < GWH_TPB / 64; w++) s += lds[w][e];
        partial[((size_t)pt * nblk + blk) * 256 + e] = s;
    }
}

/* The stencil and the FP64 Gram in one kernel: one block of GWH_FROWS waves per (point, chunk of 64 bins).  Wave
 * i evaluates operand row i -- the S stencil evaluations of component i, unrolled, projected on every detector and
 * weighted by sqrt(4 w / S_n) -- for the chunk's bins, one per lane, into LDS planes (detector, re/im) x row x bin
 * (zero for padding rows and bins past the grid); then the waves contract every plane on the FP64 matrix cores
 * (v_mfma_f64_16x16x4f64: 4 bins per instruction, lane l supplying component l % 16 of bin l / 16 as both
 * operands; wave w takes the (plane, 4-bin) slices w, w + 16, ...), and the block writes one 16x16 partial Gram,
 * summed over its waves in a fixed order.  The operand rows never leave the CU: the two-kernel path writes and
 * re-reads them through HBM (2 x 1.07 GB per 86-point pass of the 13-parameter IMRPhenomPv2 Fisher on the
 * 16225-bin 3-detector event). */
template <int MODEL, int S, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void gwh_fisher_fused_kernel(StencilArgs a,
                                                                     const EventDesc *__restrict__ events,
                                                                     int event0, double *__restrict__ partial, int nbc)
{
    __shared__ double lds[2 * GWH_MAXDET][GWH_FROWS][GWH_FUSED_BINS]; /* 64 KiB */
    const int pt = blockIdx.y, bc = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const EventDesc &E = events[event0];
    const int N = E.N;
    const int k0 = bc * GWH_FUSED_BINS + lane;
    const bool in = k0 < N;
    const int k = in ? k0 : N - 1;
    const InstCoef *C = a.coef + (size_t)pt * a.per_point;
    const double *cw = a.cw + (size_t)pt * GWH_FROWS * GWH_MAXSTEN;
    const unsigned char *nt = a.no_tof + (size_t)pt * a.per_point;
    /* wave wv evaluates rows wv, wv + WAVES, ... (unrolled: no loop for the compiler to hoist values across) */
#pragma unroll
    for (int h = 0; h < GWH_FROWS / WAVES; h++) {
        const int i = wv + WAVES * h;
        double dr[GWH_MAXDET] = {0, 0, 0, 0}, di[GWH_MAXDET] = {0, 0, 0, 0};
        if (i < a.dim) { /* uniform over the wave */
            const double f = gwh_g(E.f)[k];
            const BinPow bp = gwh_load_bin(E, k);
#pragma unroll
            for (int s = 0; s < S; s++) {
                const InstCoef &Ci = C[i * S + s];
                if (!(f <= Ci.fcut)) continue; /* h = 0 above 0.2/M of this evaluation */
                const double c = cw[i * GWH_MAXSTEN + s];
                const bool tof = !nt[i * S + s];
                if (MODEL == GWH_IMRPHENOMD) {
                    double A, phi, u, lx;
                    gwh_phenomd_fp64(Ci, f, bp, A, phi, u, lx);
                    for (int dd = 0; dd < GWH_MAXDET; dd++) {
                        if (dd >= a.d_count) break;
                        const int d = a.d_first + dd;
                        double sn, cs;
                        gwh_sincos_table(phi + (tof ? Ci.det[d][3] : 0.0) * f, sn, cs);
                        const double er = A * cs, ei = -A * sn; /* A K e^{-i(phi + 2 pi dt f)} */
                        dr[dd] += c * (er * Ci.det[d][0] - ei * Ci.det[d][1]);
                        di[dd] += c * (er * Ci.det[d][1] + ei * Ci.det[d][0]);
                    }
                } else {
                    double A, Phi, hpr, hpi, hcr, hci;
                    gwh_pv2_fp64(Ci, f, bp, A, Phi, hpr, hpi, hcr, hci);
                    for (int dd = 0; dd < GWH_MAXDET; dd++) {
                        if (dd >= a.d_count) break;
                        const int d = a.d_first + dd;
                        const double Gr = Ci.det[d][0] * hpr + Ci.det[d][1] * hcr;
                        const double Gi = Ci.det[d][0] * hpi + Ci.det[d][1] * hci;
                        double sn, cs;
                        gwh_sincos_table(Phi + (tof ? Ci.det[d][3] : 0.0) * f, sn, cs);
                        dr[dd] += c * A * (Gr * cs + Gi * sn);
                        di[dd] += c * A * (Gi * cs - Gr * sn);
                    }
                }
            }
            for (int dd = 0; dd < GWH_MAXDET; dd++) {
                if (dd >= a.d_count) break;
                const double sw = in ? sqrt(gwh_g(E.wS)[(size_t)(a.d_first + dd) * N + k]) : 0.0;
                dr[dd] *= sw;
                di[dd] *= sw;
            }
        }
        for (int dd = 0; dd < GWH_MAXDET; dd++) {
            if (dd >= a.d_count) break;
            lds[2 * dd][i][lane] = dr[dd];
            lds[2 * dd + 1][i][lane] = di[dd];
        }
    }
    __syncthreads();
    gwh_d4 acc = {0, 0, 0, 0};
    const int nslice = 2 * a.d_count * (GWH_FUSED_BINS / 4);
    for (int t = wv; t < nslice; t += WAVES) {
        const int pl = t / (GWH_FUSED_BINS / 4), q = t % (GWH_FUSED_BINS / 4);
        const double x = lds[pl][lane & 15][4 * q + (lane >> 4)];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, x, acc, 0, 0, 0);
    }
    __syncthreads(); /* the planes are read: the LDS now holds the waves' partial matrices */
    double(*part)[256] = reinterpret_cast<double(*)[256]>(&lds[0][0][0]);
    for (int j = 0; j < 4; j++) part[wv][((lane >> 4) + 4 * j) * 16 + (lane & 15)] = acc[j];
    __syncthreads();
    if (threadIdx.x < 256) {
        double sum = part[0][threadIdx.x];
        for (int w = 1; w < WAVES; w++) sum += part[w][threadIdx.x];
        partial[((size_t)pt * nbc + bc) * 256 + threadIdx.x] = sum;
    }
}

/* Fixed-order FP64 sum of the Gram partials, one block of 1024 threads per point: thread (q, e) sums element e of
 * the 16x16 partials of blocks q, q + 4, ... in increasing order, the four sums are added as (s0 + s1) + (s2 + s3),
 * and the output is symmetric (the upper and lower triangles are averaged, as calculate_fisher_elements assigns
 * out[k][j] = out[j][k]). */
__global__ __launch_bounds__(1024) void gwh_gram_reduce_kernel(const double *__restrict__ partial, int nblk, int dim,
                                                               double scale, double *__restrict__ out)
{
    __shared__ double red[4][256];
    const int pt = blockIdx.x;
    const int e = threadIdx.x & 255, q = threadIdx.x >> 8;
    const double *pp = partial + (size_t)pt * nblk * 256 + e;
    double acc = 0;
    for (int b = q; b < nblk; b += 4) acc += pp[(size_t)b * 256];
    red[q][e] = acc;
    __syncthreads();
    if (q == 0) red[0][e] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
    __syncthreads();
    if (q == 0) {
        const int i = e >> 4, j = e & 15;
        if (i < dim && j < dim) out[(size_t)pt * dim * dim + i * dim + j] = 0.5 * (red[0][e] + red[0][j * 16 + i]) * scale;
    }
}

/* FP64 LU inverse with partial pivoting per point -> marginal sigmas sqrt(Sigma_ii): one wave per point.  The
 * elimination of step k updates every element (i, j > k) once, spread over the lanes, in the same order of steps
 * as a serial elimination; the pivot is the first row of largest |a_ik| (a NaN diagonal keeps its row); lane c then
 * solves L U x = P e_c by forward and back substitution (loops of fixed length, so x stays in registers). */
__global__ __launch_bounds__(64) void gwh_fisher_invert_kernel(const double *__restrict__ F, int n, int dim,
                                                               double *__restrict__ sigma)
{
    __shared__ double a[GWH_FROWS][GWH_FROWS + 1];
    __shared__ int perm[GWH_FROWS];
    __shared__ int piv_s;
    const int pt = blockIdx.x, lane = threadIdx.x;
    if (pt >= n) return; /* the whole block */
    const double *M = F + (size_t)pt * dim * dim;
    for (int e = lane; e < dim * dim; e += 64) a[e / dim][e % dim] = M[e];
    if (lane < dim) perm[lane] = lane;
    __syncthreads();
    bool sing = false;
    for (int k = 0; k < dim; k++) {
        /* pivot: the first row i >= k of largest |a_ik| */
        if (lane == 0) {
            int piv = k;
            double mx = fabs(a[k][k]);
            for (int i = k + 1; i < dim; i++)
                if (fabs(a[i][k]) > mx) {
                    mx = fabs(a[i][k]);
                    piv = i;
                }
            piv_s = (mx == 0) ? -1 : piv;
        }
        __syncthreads();
        const int piv = piv_s;
        if (piv < 0) {
            sing = true;
            break;
        }
        if (piv != k) {
            if (lane < dim) {
                const double t = a[k][lane];
                a[k][lane] = a[piv][lane];
                a[piv][lane] = t;
            }
            if (lane == 0) {
                const int t = perm[k];
                perm[k] = perm[piv];
                perm[piv] = t;
            }
            __syncthreads();
        }
        if (lane > k && lane < dim) a[lane][k] /= a[k][k];
        __syncthreads();
        const int m = dim - k - 1;
        for (int e = lane; e < m * m; e += 64) {
            const int i = k + 1 + e / m, j = k + 1 + e % m;
            a[i][j] -= a[i][k] * a[k][j];
        }
        __syncthreads();
    }
    const int c = lane;
    if (c >= dim) return;
    if (sing) {
        sigma[(size_t)pt * dim + c] = __longlong_as_double(0x7ff8000000000000LL);
        return;
    }
    double x[GWH_FROWS];
#pragma unroll
    for (int i = 0; i < GWH_FROWS; i++) {
        x[i] = 0;
        if (i < dim) {
            double s = (perm[i] == c) ? 1.0 : 0.0;
#pragma unroll
            for (int j = 0; j < i; j++) s -= a[i][j] * x[j];
            x[i] = s;
        }
    }
#pragma unroll
    for (int i = GWH_FROWS - 1; i >= 0; i--) {
        if (i < dim) {
            double s = x[i];
#pragma unroll
            for (int j = i + 1; j < GWH_FROWS; j++)
                if (j < dim) s -= a[i][j] * x[j];
            x[i] = s / a[i][i];
        }
    }
    double xc = 0;
#pragma unroll
    for (int i = 0; i < GWH_FROWS; i++)
        if (i == c) xc = x[i];
    sigma[(size_t)pt * dim + c] = sqrt(xc);
}

/* ------------------------------------------------------------------------- */
/* host driver                                                                */
/* ------------------------------------------------------------------------- */

/* Fisher matrices of `n` points on a resident event.  d_first/d_count select the detectors whose
 * responses enter (network sum = all); ref_det is the arrival-time reference. */
static int fisher_run(int event, int model, int basis, int precision, const gwh_params *params, int n, int order,
                      double eps, int d_first, int d_count, int det_for_tof, int ref_det, double scale, double *out,
                      double *sigma, FisherScratch &F, double *raw = nullptr)
{
    Engine &E = gwh_engine();
    if (!(eps != 0) || !std::isfinite(eps))
        return gwh_set_error(GWH_ERR_INVALID, "the Fisher step must be finite and nonzero");
    Event ev;
    {
        /* copy of the descriptor; the device block stays valid because broker requests pin the
           event (gwh_event_destroy then refuses) and direct callers hold it for the call */
        std::lock_guard<std::recursive_mutex> lk(E.mu);
        ev = E.events[event];
    }
    int nmod, rc0;
    if ((rc0 = fisher_mods(model, params, n, &nmod)) != GWH_OK) return rc0;
    const bool pv2 = gwh_is_pv2(model);
    int dim = basis_dim(basis) + nmod;
    if (dim > GWH_FROWS) return gwh_set_error(GWH_ERR_UNSUPPORTED, "at most 16 Fisher parameters");
    bool sky = (basis == GWH_BASIS_PHENOMD_SKYAVG);
    if (sky && pv2)
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "Sky averaged IMRPhenomP is not supported for regular fishers.");
    int S = (order >= 4) ? 4 : 2;
    int per_point = dim * S + (sky ? 1 : 0);
    const int N = ev.N;
    const size_t nrows = (size_t)(sky ? 1 : d_count) * N * 2;
    const size_t row_bytes = GWH_FROWS * ((precision == GWH_FISHER_GRAM_FP32) ? 4 : 8);
    /* the FP64 Gram of a network Fisher runs fused with the stencil (no operand rows in HBM); the sky-averaged
       basis, the FP32 Gram and the bare derivatives go through operand rows */
    const bool fused = !raw && !sky && precision != GWH_FISHER_GRAM_FP32;
    /* points per pass: bounded operand buffer (F.max_operand, 1 GiB unless a caller's budget says less); fused,
       the partial Gram matrices take its place */
    size_t per_pt_bytes = nrows * row_bytes;
    const int nplanes = (int)(nrows / N); /* (detector, re/im) planes of a point's operand rows */
    const int blk_per_plane = (N + GWH_GRAM_ROWS - 1) / GWH_GRAM_ROWS;
    int nblk = fused ? (N + GWH_FUSED_BINS - 1) / GWH_FUSED_BINS : nplanes * blk_per_plane;
    if (fused) per_pt_bytes = (size_t)nblk * 256 * 8;
    int chunk = (int)std::max<size_t>(1, F.max_operand / per_pt_bytes);
    if (chunk > n) chunk = n;
    if (chunk > 65535) chunk = 65535;
    if (fused && chunk > GWH_FUSED_MAX_POINTS) chunk = GWH_FUSED_MAX_POINTS;
    int rc;
    if (!fused && (rc = F.deriv.ensure((size_t)chunk * per_pt_bytes)) != GWH_OK) return rc;
    if ((rc = F.gram.ensure((size_t)chunk * nblk * 256 * 8)) != GWH_OK) return rc;
    if ((rc = F.out.ensure((size_t)n * dim * dim * 8)) != GWH_OK) return rc;
    std::vector<gwh_params> inst;
    std::vector<unsigned char> no_tof;
    std::vector<double> cw;
    int slot = 0;
    for (int p0 = 0; p0 < n; p0 += chunk, slot ^= 1) {
        int np = std::min(chunk, n - p0);
        inst.clear();
        no_tof.clear();
        cw.assign((size_t)np * GWH_FROWS * GWH_MAXSTEN, 0.0);
        inst.reserve((size_t)np * per_point);
        for (int q = 0; q < np; q++) {
            PointStencil ps;
            build_stencil(params[p0 + q], basis, gwh_ppe_variant(model), nmod, order, eps, det_for_tof, ref_det,
                          inst, no_tof, ps);
            if (sky) {
                inst.push_back(params[p0 + q]); /* central amplitude: the caller's parameters */
                no_tof.push_back(0);
            }
            for (int i = 0; i < GWH_FROWS; i++)
                for (int s = 0; s < GWH_MAXSTEN; s++) cw[((size_t)q * GWH_FROWS + i) * GWH_MAXSTEN + s] = ps.c[i][s];
        }
        int ni = (int)inst.size();
        if ((rc = F.params.ensure((size_t)ni * sizeof(gwh_params))) != GWH_OK) return rc;
        if ((rc = F.coef.ensure((size_t)ni * sizeof(InstCoef))) != GWH_OK) return rc;
        if ((rc = F.misc.ensure(cw.size() * 8 + no_tof.size() + 256)) != GWH_OK) return rc;
        /* stage the pass in pinned memory (asynchronous copies; the device buffers are reused in stream order) */
        const size_t rec_bytes = (size_t)ni * sizeof(gwh_params), stage_bytes = rec_bytes + cw.size() * 8 + no_tof.size();
        if (F.stage_done[slot]) GWH_HIP(hipEventSynchronize(F.stage_done[slot])); /* its last copy has run */
        else GWH_HIP(hipEventCreateWithFlags(&F.stage_done[slot], hipEventDisableTiming));
        if (F.stage_cap[slot] < stage_bytes) {
            if (F.stage[slot]) GWH_HIP(hipHostFree(F.stage[slot]));
            F.stage[slot] = nullptr;
            F.stage_cap[slot] = 0;
            GWH_HIP(hipHostMalloc(&F.stage[slot], stage_bytes + stage_bytes / 4, hipHostMallocDefault));
            F.stage_cap[slot] = stage_bytes + stage_bytes / 4;
        }
        char *st = (char *)F.stage[slot];
        memcpy(st, inst.data(), rec_bytes);
        memcpy(st + rec_bytes, cw.data(), cw.size() * 8);
        memcpy(st + rec_bytes + cw.size() * 8, no_tof.data(), no_tof.size());
        GWH_HIP(hipMemcpyAsync(F.params.p, st, rec_bytes, hipMemcpyHostToDevice, F.stream));
        char *misc = (char *)F.misc.p;
        GWH_HIP(hipMemcpyAsync(misc, st + rec_bytes, cw.size() * 8 + no_tof.size(), hipMemcpyHostToDevice, F.stream));
        GWH_HIP(hipEventRecord(F.stage_done[slot], F.stream));
        if ((rc = gwh_setup_launch(nullptr, event, model, F.params.as<const gwh_params>(), ni, F.coef.as<InstCoef>(),
                                   F.stream)) != GWH_OK)
            return rc;
        StencilArgs a;
        a.coef = F.coef.as<const InstCoef>();
        a.cw = (const double *)misc;
        a.no_tof = (const unsigned char *)(misc + cw.size() * 8);
        a.dim = dim;
        a.S = S;
        a.per_point = per_point;
        a.d_first = d_first;
        a.d_count = sky ? 1 : d_count;
        a.skyavg = sky;
        a.sa_sign = (S == 4) ? -1.0 : 1.0;
        a.raw = nullptr;
        dim3 g((N + GWH_TPB - 1) / GWH_TPB, np, GWH_FROWS); /* z: operand row */
        const EventDesc *evp = (const EventDesc *)E.d_events;
        if (raw) {
            /* bare derivatives of one point on one detector (calculate_derivatives) */
            if ((rc = F.raw.ensure((size_t)dim * N * 16)) != GWH_OK) return rc;
            a.raw = F.raw.as<double>();
            a.d_count = 1;
            dim3 g1((N + GWH_TPB - 1) / GWH_TPB, 1, dim);
            launch_stencil<double>(pv2, g1, F.stream, a, evp, event, F.deriv.as<double>());
            GWH_HIP(hipGetLastError());
            GWH_HIP(hipMemcpyAsync(raw, F.raw.p, (size_t)dim * N * 16, hipMemcpyDeviceToHost, F.stream));
            GWH_HIP(hipStreamSynchronize(F.stream));
            return GWH_OK;
        }
        if (fused) {
            const dim3 gf(nblk, np), bf(GWH_FUSED_WAVES * 64);
            if (!pv2) {
                if (S == 4)
                    hipLaunchKernelGGL((gwh_fisher_fused_kernel<GWH_IMRPHENOMD, 4, GWH_FUSED_WAVES>), gf, bf, 0, F.stream, a, evp, event,
                                       F.gram.as<double>(), nblk);
                else
                    hipLaunchKernelGGL((gwh_fisher_fused_kernel<GWH_IMRPHENOMD, 2, GWH_FUSED_WAVES>), gf, bf, 0, F.stream, a, evp, event,
                                       F.gram.as<double>(), nblk);
            } else {
                if (S == 4)
                    hipLaunchKernelGGL((gwh_fisher_fused_kernel<GWH_IMRPHENOMPV2, 4, GWH_FUSED_WAVES>), gf, bf, 0, F.stream, a, evp,
                                       event, F.gram.as<double>(), nblk);
                else
                    hipLaunchKernelGGL((gwh_fisher_fused_kernel<GWH_IMRPHENOMPV2, 2, GWH_FUSED_WAVES>), gf, bf, 0, F.stream, a, evp,
                                       event, F.gram.as<double>(), nblk);
            }
        } else if (precision != GWH_FISHER_GRAM_FP32) {
            launch_stencil<double>(pv2, g, F.stream, a, evp, event, F.deriv.as<double>());
            GWH_HIP(hipGetLastError());
            hipLaunchKernelGGL(gwh_gram_kernel<double>, dim3(nblk, np), dim3(GWH_TPB), 0, F.stream,
                               F.deriv.as<const double>(), N, nplanes, blk_per_plane, F.gram.as<double>(), nblk);
        } else {
            launch_stencil<float>(pv2, g, F.stream, a, evp, event, F.deriv.as<float>());
            GWH_HIP(hipGetLastError());
            hipLaunchKernelGGL(gwh_gram_kernel<float>, dim3(nblk, np), dim3(GWH_TPB), 0, F.stream,
                               F.deriv.as<const float>(), N, nplanes, blk_per_plane, F.gram.as<double>(), nblk);
        }
        GWH_HIP(hipGetLastError());
        hipLaunchKernelGGL(gwh_gram_reduce_kernel, dim3(np), dim3(1024), 0, F.stream, F.gram.as<const double>(),
                           nblk, dim, scale, F.out.as<double>() + (size_t)p0 * dim * dim);
        GWH_HIP(hipGetLastError());
    }
    if (sigma) {
        if ((rc = F.sig.ensure((size_t)n * dim * 8)) != GWH_OK) return rc;
        hipLaunchKernelGGL(gwh_fisher_invert_kernel, dim3(n), dim3(64), 0, F.stream,
                           F.out.as<const double>(), n, dim, F.sig.as<double>());
        GWH_HIP(hipGetLastError());
        GWH_HIP(hipMemcpyAsync(sigma, F.sig.p, (size_t)n * dim * 8, hipMemcpyDeviceToHost, F.stream));
    }
    GWH_HIP(hipMemcpyAsync(out, F.out.p, (size_t)n * dim * dim * 8, hipMemcpyDeviceToHost, F.stream));
    GWH_HIP(hipStreamSynchronize(F.stream));
    return GWH_OK;
}

int gwh_fisher_batch_with(int event, int model, int basis, int precision, const gwh_params *params, int n, int dim,
                          double epsilon, double *out, double *sigma, FisherScratch &F)
{
    Engine &E = gwh_engine();
    int rc = gwh_ensure_init();
    if (rc != GWH_OK) return rc;
    if (n <= 0) return GWH_OK;
    if (!params || !out) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    Event ev;
    {
        std::lock_guard<std::recursive_mutex> lk(E.mu);
        if (event < 0 || event >= GWH_MAX_EVENTS || !E.events[event].live)
            return gwh_set_error(GWH_ERR_INVALID, "unknown event");
        ev = E.events[event];
    }
    int nmod;
    if ((rc = fisher_mods(model, params, n, &nmod)) != GWH_OK) return rc;
    if (basis_dim(basis) < 0 || basis_dim(basis) + nmod != dim)
        return gwh_set_error(GWH_ERR_INVALID, "dimension does not match the parameter basis (and ppE terms)");
    if (precision != GWH_FP64 && precision != GWH_MIXED && precision != GWH_FISHER_GRAM_FP32)
        return gwh_set_error(GWH_ERR_INVALID, "bad precision");
    if ((basis == GWH_BASIS_PV2_FULL || basis == GWH_BASIS_PV2_MCMC) != gwh_is_pv2(model))
        return gwh_set_error(GWH_ERR_INVALID, "basis does not belong to the model");
    /* the event's weights use the likelihood's df = f[N/2]-f[N/2-1]; the Fisher uses f[1]-f[0] */
    double scale = ev.df_fisher / ev.df_like;
    return fisher_run(event, model, basis, precision, params, n, 2, epsilon, 0, ev.D, ev.det[0], ev.det[0], scale, out,
                      sigma, F);
}

size_t gwh_fisher_batch_bytes(int event, int model, int basis, int precision, const gwh_params *params, int n,
                              size_t max_operand, const FisherScratch &F, bool want_sigma)
{
    Engine &E = gwh_engine();
    Event ev;
    {
        std::lock_guard<std::recursive_mutex> lk(E.mu);
        if (event < 0 || event >= GWH_MAX_EVENTS || !E.events[event].live) return 0;
        ev = E.events[event];
    }
    int nmod;
    if (n <= 0 || !params || fisher_mods(model, params, n, &nmod) != GWH_OK || basis_dim(basis) < 0) return 0;
    /* the buffers fisher_run ensures (order 2) */
    const int dim = basis_dim(basis) + nmod;
    const bool sky = (basis == GWH_BASIS_PHENOMD_SKYAVG);
    const size_t per_point = (size_t)dim * 2 + (sky ? 1 : 0);
    const size_t nrows = (size_t)(sky ? 1 : ev.D) * ev.N * 2;
    const bool fused = !sky && precision != GWH_FISHER_GRAM_FP32;
    const size_t nblk = fused ? (ev.N + GWH_FUSED_BINS - 1) / GWH_FUSED_BINS
                              : (nrows / ev.N) * ((ev.N + GWH_GRAM_ROWS - 1) / GWH_GRAM_ROWS);
    const size_t per_pt_bytes = fused ? nblk * 256 * 8 : nrows * GWH_FROWS * ((precision == GWH_FISHER_GRAM_FP32) ? 4 : 8);
    size_t chunk = std::max<size_t>(1, max_operand / per_pt_bytes);
    chunk = std::min<size_t>(chunk, std::min(n, 65535));
    if (fused) chunk = std::min<size_t>(chunk, GWH_FUSED_MAX_POINTS);
    const size_t ni = chunk * per_point;
    return (fused ? F.deriv.cap : gwh_grown_capacity(chunk * per_pt_bytes, F.deriv.cap)) +
           gwh_grown_capacity(chunk * nblk * 256 * 8, F.gram.cap) +
           gwh_grown_capacity((size_t)n * dim * dim * 8, F.out.cap) +
           (want_sigma ? gwh_grown_capacity((size_t)n * dim * 8, F.sig.cap) : F.sig.cap) +
           gwh_grown_capacity(ni * sizeof(gwh_params), F.params.cap) +
           gwh_grown_capacity(ni * sizeof(InstCoef), F.coef.cap) +
           gwh_grown_capacity(chunk * GWH_FROWS * GWH_MAXSTEN * 8 + ni + 256, F.misc.cap) + F.raw.cap;
}

extern "C" {

/* detector_util.cpp:573-636 name matching (exact strings only; anything else is unsupported) */
int gwh_detector_id(const char *name)
{
    if (!name) return -1;
    static const struct {
        const char *s;
        int id;
    } names[] = {{"Hanford", GWH_HANFORD},
                 {"hanford", GWH_HANFORD},
                 {"Livingston", GWH_LIVINGSTON},
                 {"livingston", GWH_LIVINGSTON},
                 {"Virgo", GWH_VIRGO},
                 {"virgo", GWH_VIRGO},
                 {"Kagra", GWH_KAGRA},
                 {"kagra", GWH_KAGRA},
                 {"Indigo", GWH_INDIGO},
                 {"indigo", GWH_INDIGO},
                 {"Cosmic Explorer", GWH_CE},
                 {"cosmic explorer", GWH_CE},
                 {"CE", GWH_CE},
                 {"Einstein Telescope 1", GWH_ET1},
                 {"einstein telescope 1", GWH_ET1},
                 {"ET1", GWH_ET1},
                 {"Einstein Telescope 2", GWH_ET2},
                 {"einstein telescope 2", GWH_ET2},
                 {"ET2", GWH_ET2},
                 {"Einstein Telescope 3", GWH_ET3},
                 {"einstein telescope 3", GWH_ET3},
                 {"ET3", GWH_ET3}};
    for (auto &n : names)
        if (!strcmp(n.s, name)) return n.id;
    return -1;
}

double gwh_dtoa(double RA, double DEC, double gmst, int detector1, int detector2)
{
    if (detector1 < 0 || detector1 >= GWH_NDET_TABLE || detector2 < 0 || detector2 >= GWH_NDET_TABLE) {
        gwh_set_error(GWH_ERR_UNSUPPORTED, "unsupported detector");
        return __builtin_nan("");
    }
    return host_dtoa(RA, DEC, gmst, detector1, detector2);
}

/* util.cpp:1797-1846: GPS seconds -> Greenwich mean sidereal time in radians (the reference's
 * low-order series in Julian days from J2000, hours wrapped to [0, 24)) */
double gwh_gps_to_gmst_radian(double gps)
{
    const double jd_j2000 = 2451545.0, gps_at_j2000 = 630763213.0;
    double jd = jd_j2000 + (gps - gps_at_j2000) / 86400.0;
    double jd0 = floor(jd) + 0.5;
    double hours_ut = (jd - jd0) * 24.0;
    double d0 = jd0 - jd_j2000, centuries = (jd - jd_j2000) / 36525.0;
    double g = 6.697374558 + 0.06570982441908 * d0 + 1.00273790935 * hours_ut + 0.000026 * centuries * centuries;
    double whole = (double)((long long)floor(g) % 24);
    return (whole + (g - floor(g))) * M_PI / 12.0;
}

int gwh_unpack_parameters(const gwh_params *params, int basis, double *v, int *log_factors)
{
    if (!params || !v) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    int dim = basis_dim(basis);
    if (dim < 0) return gwh_set_error(GWH_ERR_INVALID, "unknown basis");
    int lf[GWH_FROWS];
    unpack(*params, basis, v, lf);
    if (log_factors)
        for (int i = 0; i < dim; i++) log_factors[i] = lf[i];
    return GWH_OK;
}

int gwh_repack_parameters(const double *v, int basis, gwh_params *params)
{
    if (!params || !v) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    if (basis_dim(basis) < 0) return gwh_set_error(GWH_ERR_INVALID, "unknown basis");
    repack(v, basis, *params);
    return GWH_OK;
}

/* one-point Fisher on a temporary event holding the PSD for (reference detector, detector): the
 * time-of-flight phase of the second is relative to the first; the weights use df = f[1]-f[0]
 * (fisher.cpp:2746).  deriv != null: return the bare derivatives instead (calculate_derivatives). */
static int single_point(const double *frequencies, int length, int model, int basis, int detector,
                        int reference_detector, double *out, int dim, const gwh_params *params, int order,
                        const double *psd, double epsilon, double *deriv)
{
    Engine &E = gwh_engine();
    int rc = gwh_ensure_init();
    if (rc != GWH_OK) return rc;
    if (!frequencies || !(out || deriv) || !params || length < 2 || (!psd && !deriv))
        return gwh_set_error(GWH_ERR_INVALID, "null argument or length < 2");
    int nmod;
    if ((rc = fisher_mods(model, params, 1, &nmod)) != GWH_OK) return rc;
    if (basis_dim(basis) < 0 || basis_dim(basis) + nmod != dim)
        return gwh_set_error(GWH_ERR_INVALID, "dimension does not match the parameter basis (and ppE terms)");
    if ((basis == GWH_BASIS_PV2_FULL || basis == GWH_BASIS_PV2_MCMC) != gwh_is_pv2(model))
        return gwh_set_error(GWH_ERR_INVALID, "basis does not belong to the model");
    if (order != 2 && order != 4) return gwh_set_error(GWH_ERR_INVALID, "order must be 2 or 4");
    if (detector < 0 || detector >= GWH_NDET_TABLE || reference_detector < 0 || reference_detector >= GWH_NDET_TABLE)
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "unsupported detector");
    if (basis == GWH_BASIS_PHENOMD_SKYAVG && !params->sky_average)
        return gwh_set_error(GWH_ERR_INVALID, "the sky-averaged basis needs params->sky_average");
    if (basis != GWH_BASIS_PHENOMD_SKYAVG && params->sky_average)
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "sky_average selects the sky-averaged IMRPhenomD basis only");
    int dets[2] = {reference_detector, detector};
    int nd = (detector != reference_detector) ? 2 : 1;
    if (nd == 1) dets[0] = detector;
    /* the PSD of the Fisher detector on both rows (the reference detector's row only carries the arrival-time
     * reference); derivatives alone need no PSD: a unit one */
    std::vector<double> unit;
    const double *ps = psd;
    if (!ps) {
        unit.assign((size_t)length, 1.0);
        ps = unit.data();
    }
    const double *fp[2] = {frequencies, frequencies}, *pp[2] = {ps, ps};
    int ev;
    /* a cached event (no allocation after the first call on these arrays) */
    if ((rc = gwh_single_call_event(fp, length, nd, dets, nullptr, pp, GWH_SIMPSONS, nullptr, 0, &ev, 1)) != GWH_OK)
        return rc;
    E.events[ev].pins++;
    rc = fisher_run(ev, model, basis, GWH_FP64, params, 1, order, epsilon, nd - 1, 1, detector, reference_detector,
                    1.0, out, nullptr, E.fisher, deriv);
    E.events[ev].pins--;
    return rc;
}

static int gr_model_of(int basis)
{
    return (basis == GWH_BASIS_PV2_FULL || basis == GWH_BASIS_PV2_MCMC) ? GWH_IMRPHENOMPV2 : GWH_IMRPHENOMD;
}

int gwh_fisher_derivatives_model(const double *frequencies, int length, int model, int basis, int detector,
                                 int reference_detector, int dim, const gwh_params *params, int order, double epsilon,
                                 double *deriv)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (!deriv) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    return single_point(frequencies, length, model, basis, detector, reference_detector, nullptr, dim, params, order,
                        nullptr, epsilon, deriv);
}

int gwh_fisher_numerical_model(const double *frequencies, int length, int model, int basis, int detector,
                               int reference_detector, double *out, int dim, const gwh_params *params, int order,
                               const double *psd, double epsilon)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (!out || !psd) return gwh_set_error(GWH_ERR_INVALID, "null argument");
    return single_point(frequencies, length, model, basis, detector, reference_detector, out, dim, params, order, psd,
                        epsilon, nullptr);
}

int gwh_fisher_derivatives(const double *frequencies, int length, int basis, int detector, int reference_detector,
                           int dim, const gwh_params *params, int order, double epsilon, double *deriv)
{
    return gwh_fisher_derivatives_model(frequencies, length, gr_model_of(basis), basis, detector, reference_detector,
                                        dim, params, order, epsilon, deriv);
}

int gwh_fisher_numerical(const double *frequencies, int length, int basis, int detector, int reference_detector,
                         double *out, int dim, const gwh_params *params, int order, const double *psd, double epsilon)
{
    return gwh_fisher_numerical_model(frequencies, length, gr_model_of(basis), basis, detector, reference_detector,
                                      out, dim, params, order, psd, epsilon);
}

int gwh_fisher_batch(int event, int model, int basis, int precision, const gwh_params *params, int n, int dim,
                     double epsilon, double *out, double *sigma)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int rc = gwh_ensure_init();
    if (rc != GWH_OK) return rc;
    return gwh_fisher_batch_with(event, model, basis, precision, params, n, dim, epsilon, out, sigma, E.fisher);
}

} /* extern "C" */
