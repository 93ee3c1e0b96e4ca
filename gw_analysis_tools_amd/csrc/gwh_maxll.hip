/*
 * gwh_maxll.hip -- likelihoods maximised over the coalescence time and phase (and the distance; the
 * unaligned statistic also over the polarisation) through FFTs of the matched-filter integrand.
 *
 *   fill     one lane per (bin, instance): the instance's FP64 waveform, g(f) = 4 d*(f) h(f) / S(f) per
 *            detector (and polarisation) into a batch of transform inputs, and fixed-order block partials
 *            of the Simpson sums (h|h) (unaligned: (h+|h+), (hx|hx), Re(h+|hx))
 *   rocFFT   batched in-place complex forward transforms of length N (FFTW_FORWARD, as the reference)
 *   reduce   one block per instance: per detector the maximum over the transform (first index of the
 *            maximum on ties), the statistic, and the network sum in detector order
 *
 * Reference: src/mcmc_gw.cpp:493-556 (maximized_coal_Log_Likelihood_internal), :595-654
 * (maximized_Log_Likelihood_aligned_spin_internal), :660-800 (..._unaligned_spin_internal), :2607-2700 (the
 * intrinsic MCMC likelihood: extrinsic parameters fixed, the statistic summed over detectors).
 */
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gwh_engine.h"

#define GWH_MX_TPB 256
#define GWH_MX_BYTES (1ull << 30) /* transform inputs per pass */

/* ------------------------------------------------------------------------- */
/* kernels                                                                    */
/* ------------------------------------------------------------------------- */

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

/* sums NV per-thread values over the block in a fixed order; the block's totals land in out[0..NV) */
template <int NV> __device__ __forceinline__ void block_sum_store(double (&v)[NV], double *out)
{
    __shared__ double lds[GWH_MX_TPB / 64][NV > 0 ? NV : 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) lds[wv][q] = v[q];
    __syncthreads();
    if ((int)threadIdx.x < NV) {
        double s = lds[0][threadIdx.x];
        for (int w = 1; w < GWH_MX_TPB / 64; w++) s += lds[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

/* Transform inputs of the instances [i0, i0 + gridDim.y) of one resident event: buf[((j D + d) P + p) N + k],
 * p = 0 plus (IMRPhenomD: the plus polarisation), 1 cross (IMRPhenomPv2).  Partials per (j, d, block):
 * sum wS |h+|^2, sum wS |hx|^2, sum wS Re(h+ hx*), wS the event's Simpson weight x 4 / S. */
template <int MODEL>
__global__ __launch_bounds__(GWH_MX_TPB) void gwh_maxll_fill_kernel(const EventDesc *__restrict__ events, int ev,
                                                                    const InstCoef *__restrict__ coefs, double dfac,
                                                                    double2 *__restrict__ buf,
                                                                    double *__restrict__ part)
{
    const EventDesc &E = events[ev];
    const int N = E.N, D = E.D;
    const int P = (MODEL == GWH_IMRPHENOMPV2) ? 2 : 1;
    const int k = blockIdx.x * GWH_MX_TPB + threadIdx.x;
    const int j = blockIdx.y, nblk = gridDim.x;
    const InstCoef &C = coefs[j];
    double2 hp = make_double2(0, 0), hc = make_double2(0, 0);
    if (k < N) {
        const double f = gwh_g(E.f)[k];
        if (f <= C.fcut) {
            const BinPow bp = gwh_load_bin(E, k);
            if (MODEL == GWH_IMRPHENOMD) {
                double A, phi, u, lx;
                gwh_phenomd_fp64(C, f, bp, A, phi, u, lx);
                double s, c;
                gwh_sincos_table(phi, s, c);
                hp = make_double2(A * c * C.wf_a, -A * s * C.wf_a); /* h+ = A e^{-i phi} (1 + cos^2 i) / 2 */
            } else {
                double A, Phi, hpr, hpi, hcr, hci;
                gwh_pv2_fp64(C, f, bp, A, Phi, hpr, hpi, hcr, hci);
                double s, c;
                gwh_sincos_table(Phi, s, c);
                const double pr = A * (hpr * c + hpi * s), pi = A * (hpi * c - hpr * s);
                const double xr = A * (hcr * c + hci * s), xi = A * (hci * c - hcr * s);
                const double c2z = C.wf_a, s2z = C.wf_b;
                hp = make_double2(c2z * pr + s2z * xr, c2z * pi + s2z * xi);
                hc = make_double2(c2z * xr - s2z * pr, c2z * xi - s2z * pi);
            }
        }
    }
    for (int d = 0; d < D; d++) {
        double v[3] = {0, 0, 0};
        if (k < N) {
            const double ws = gwh_g(E.wS)[(size_t)d * N + k];
            const double2 x = gwh_ld2(E.dwS, (size_t)d * N + k); /* d ws */
            const double g = 1.0 / (dfac * simpson_w(k, N));    /* ws / (4 / S) */
            const double2 cd = make_double2(x.x * g, -x.y * g); /* 4 d* / S */
            const size_t o = ((size_t)(j * D + d) * P) * N + k;
            buf[o] = cmul(cd, hp);
            if (P == 2) buf[o + N] = cmul(cd, hc);
            v[0] = ws * (hp.x * hp.x + hp.y * hp.y);
            v[1] = ws * (hc.x * hc.x + hc.y * hc.y);
            v[2] = ws * (hp.x * hc.x + hp.y * hc.y);
        }
        block_sum_store<3>(v, part + ((size_t)(j * D + d) * nblk + blockIdx.x) * 3);
    }
}

/* Transform inputs from caller arrays (one detector): g = 4 d* h / S; partials of sum w_k |h|^2 / S etc.
 * with the Simpson index weights w_k (the step and the 4 are applied by the reduction). */
__global__ __launch_bounds__(GWH_MX_TPB) void gwh_maxll_fill_given_kernel(int N, const double2 *__restrict__ data,
                                                                          const double *__restrict__ psd,
                                                                          const double2 *__restrict__ hpl,
                                                                          const double2 *__restrict__ hcr,
                                                                          double2 *__restrict__ buf,
                                                                          double *__restrict__ part)
{
    const int k = blockIdx.x * GWH_MX_TPB + threadIdx.x;
    double v[3] = {0, 0, 0};
    if (k < N) {
        const double2 x = data[k], hp = hpl[k];
        const double is = 1.0 / psd[k];
        const double2 cd = make_double2(4.0 * x.x * is, -4.0 * x.y * is);
        buf[k] = cmul(cd, hp);
        const double w = simpson_w(k, N) * is;
        v[0] = w * (hp.x * hp.x + hp.y * hp.y);
        if (hcr) {
            const double2 hc = hcr[k];
            buf[(size_t)N + k] = cmul(cd, hc);
            v[1] = w * (hc.x * hc.x + hc.y * hc.y);
            v[2] = w * (hp.x * hc.x + hp.y * hc.y);
        }
    }
    block_sum_store<3>(v, part + (size_t)blockIdx.x * 3);
}

/* (value, index) maximum over the block; ties keep the smaller index (the first maximum) */
__device__ __forceinline__ void block_argmax(double &v, int &idx)
{
    __shared__ double sv[GWH_MX_TPB];
    __shared__ int si[GWH_MX_TPB];
    sv[threadIdx.x] = v;
    si[threadIdx.x] = idx;
    __syncthreads();
    for (int w = GWH_MX_TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double a = sv[threadIdx.x], b = sv[threadIdx.x + w];
            const int ia = si[threadIdx.x], ib = si[threadIdx.x + w];
            if (b > a || (b == a && ib < ia) || (a != a && b == b)) {
                sv[threadIdx.x] = b;
                si[threadIdx.x] = ib;
            }
        }
        __syncthreads();
    }
    v = sv[0];
    idx = si[0];
    __syncthreads();
}

/* mode 0 aligned (mcmc_gw.cpp:604-654), 1 unaligned (:669-800), 2 coalescence (:502-556; out3 = {logL, tc,
 * phiRef}).  One block per instance; hh_scale converts the partial sums to 4 x Simpson sums with the
 * statistic's step; df is that step. */
__global__ __launch_bounds__(GWH_MX_TPB) void gwh_maxll_reduce_kernel(int N, int D, int nblk, int mode,
                                                                      const double2 *__restrict__ buf,
                                                                      const double *__restrict__ part,
                                                                      double hh_scale, double df,
                                                                      double *__restrict__ out)
{
    __shared__ double sums[3];
    const int j = blockIdx.x;
    const int P = (mode == 1) ? 2 : 1;
    double total = 0;
    for (int d = 0; d < D; d++) {
        const size_t pd = (size_t)(j * D + d) * nblk;
        if ((int)threadIdx.x < 3) {
            double s = 0;
            for (int b = 0; b < nblk; b++) s += part[(pd + b) * 3 + threadIdx.x];
            sums[threadIdx.x] = s * hh_scale;
        }
        __syncthreads();
        const double2 *o = buf + (size_t)(j * D + d) * P * N;
        double best = -INFINITY;
        int bi = N;
        if (mode == 0) {
            for (int k = threadIdx.x; k < N; k += GWH_MX_TPB) {
                const double2 z = o[k];
                const double v = z.x * z.x + z.y * z.y;
                if (v > best) best = v, bi = k;
            }
        } else if (mode == 2) {
            for (int k = threadIdx.x; k < N; k += GWH_MX_TPB) {
                const double2 z = o[k];
                const double v = hypot(z.x, z.y);
                if (v > best) best = v, bi = k;
            }
        } else {
            const double hp = sqrt(sums[0]), hc = sqrt(sums[1]);
            const double ipc = sums[2] / (hp * hc);
            for (int k = threadIdx.x; k < N; k += GWH_MX_TPB) {
                const double2 a = o[k], b = o[(size_t)N + k];
                const double2 rp = make_double2(df * a.x / hp, df * a.y / hp);
                const double2 rc = make_double2(df * b.x / hc, df * b.y / hc);
                const double rp2 = rp.x * rp.x + rp.y * rp.y, rc2 = rc.x * rc.x + rc.y * rc.y;
                const double gam = rp.x * rc.x + rp.y * rc.y;
                const double lam = (rp2 + rc2 - 2 * gam * ipc +
                                    sqrt((rp2 - rc2) * (rp2 - rc2) + 4. * (ipc * rp2 - gam) * (ipc * rc2 - gam))) /
                                   (1. - ipc * ipc);
                if (lam > best) best = lam, bi = k;
            }
        }
        block_argmax(best, bi);
        if (threadIdx.x == 0) {
            if (mode == 0) {
                total += .5 * (best * df * df) / sums[0];
            } else if (mode == 1) {
                total += .25 * best;
            } else if (bi < N) {
                const double2 z = o[bi];
                out[3 * j] = -0.5 * (sums[0] - 2 * best * df);
                out[3 * j + 1] = (double)bi / N * (1. / df);
                out[3 * j + 2] = atan2(z.y, z.x);
            } else { /* no finite maximum */
                const double nan = __longlong_as_double(0x7ff8000000000000LL);
                out[3 * j] = out[3 * j + 1] = out[3 * j + 2] = nan;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && mode != 2) out[j] = total;
}

/* ------------------------------------------------------------------------- */
/* rocFFT plans                                                               */
/* ------------------------------------------------------------------------- */

namespace {

struct FftPlan {
    size_t n = 0, batch = 0;
    rocfft_plan plan = nullptr;
    size_t work = 0;
    uint64_t used = 0;
};
std::vector<FftPlan> g_plans; /* engine lock held */
bool g_rocfft = false;
uint64_t g_plan_clock = 0;

int rf(rocfft_status s, const char *what)
{
    if (s != rocfft_status_success)
        return gwh_set_error(GWH_ERR_HIP, std::string("rocFFT: ") + what + " failed (" + std::to_string((int)s) + ")");
    return GWH_OK;
}

/* in-place forward transforms of `batch` contiguous complex FP64 sequences of length n on `st` */
int fft_forward(double2 *buf, size_t n, size_t batch, hipStream_t st)
{
    Engine &E = gwh_engine();
    int rc;
    if (!g_rocfft) {
        if ((rc = rf(rocfft_setup(), "setup")) != GWH_OK) return rc;
        g_rocfft = true;
    }
    FftPlan *p = nullptr;
    for (auto &q : g_plans)
        if (q.n == n && q.batch == batch) p = &q;
    if (!p) {
        if (g_plans.size() >= 8) { /* keep a few plans: drop the least recently used */
            size_t old = 0;
            for (size_t i = 1; i < g_plans.size(); i++)
                if (g_plans[i].used < g_plans[old].used) old = i;
            GWH_HIP(hipStreamSynchronize(st));
            rocfft_plan_destroy(g_plans[old].plan);
            g_plans.erase(g_plans.begin() + old);
        }
        FftPlan q;
        q.n = n;
        q.batch = batch;
        const size_t len[1] = {n};
        if ((rc = rf(rocfft_plan_create(&q.plan, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                                        rocfft_precision_double, 1, len, batch, nullptr),
                     "plan_create")) != GWH_OK)
            return rc;
        if ((rc = rf(rocfft_plan_get_work_buffer_size(q.plan, &q.work), "work size")) != GWH_OK) return rc;
        g_plans.push_back(q);
        p = &g_plans.back();
    }
    p->used = ++g_plan_clock;
    if (p->work && (rc = E.fftwork.ensure(p->work)) != GWH_OK) return rc;
    rocfft_execution_info info = nullptr;
    if ((rc = rf(rocfft_execution_info_create(&info), "execution_info")) != GWH_OK) return rc;
    rc = rf(rocfft_execution_info_set_stream(info, st), "set_stream");
    if (rc == GWH_OK && p->work) rc = rf(rocfft_execution_info_set_work_buffer(info, E.fftwork.p, p->work), "work buffer");
    void *io[1] = {buf};
    if (rc == GWH_OK) rc = rf(rocfft_execute(p->plan, io, nullptr, info), "execute");
    rocfft_execution_info_destroy(info);
    return rc;
}

double nan_with(int code, const std::string &msg)
{
    gwh_set_error(code, msg);
    return std::nan("");
}

/* the single-call statistics on caller arrays (one detector) */
double given(const double *data, const double *psd, const double *f, const double *hplus, const double *hcross,
             size_t length, int mode, double *tc, double *phiRef)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    if (gwh_ensure_init() != GWH_OK) return std::nan("");
    if (!data || !psd || !f || !hplus || (mode == 1 && !hcross)) return nan_with(GWH_ERR_INVALID, "null argument");
    if (length < 2 || length > (1u << 30)) return nan_with(GWH_ERR_INVALID, "length must be 2 .. 2^30");
    const int N = (int)length, P = (mode == 1) ? 2 : 1;
    const int nblk = (N + GWH_MX_TPB - 1) / GWH_MX_TPB;
    if (E.data.ensure((size_t)N * 16) != GWH_OK || E.psd.ensure((size_t)N * 8) != GWH_OK ||
        E.out1.ensure((size_t)N * 16) != GWH_OK || E.out2.ensure((size_t)N * 16) != GWH_OK ||
        E.fft.ensure((size_t)P * N * 16) != GWH_OK || E.mxpart.ensure((size_t)nblk * 3 * 8) != GWH_OK ||
        E.mxres.ensure(3 * 8) != GWH_OK)
        return std::nan("");
    hipStream_t st = E.stream;
    auto up = [&](void *dst, const void *src, size_t b) {
        return hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    if (!up(E.data.p, data, (size_t)N * 16) || !up(E.psd.p, psd, (size_t)N * 8) ||
        !up(E.out1.p, hplus, (size_t)N * 16) || (mode == 1 && !up(E.out2.p, hcross, (size_t)N * 16)))
        return nan_with(GWH_ERR_HIP, "upload failed");
    hipLaunchKernelGGL(gwh_maxll_fill_given_kernel, dim3(nblk), dim3(GWH_MX_TPB), 0, st, N,
                       E.data.as<const double2>(), E.psd.as<const double>(), E.out1.as<const double2>(),
                       mode == 1 ? E.out2.as<const double2>() : nullptr, E.fft.as<double2>(), E.mxpart.as<double>());
    if (hipGetLastError() != hipSuccess) return nan_with(GWH_ERR_HIP, "fill launch failed");
    if (fft_forward(E.fft.as<double2>(), (size_t)N, P, st) != GWH_OK) return std::nan("");
    /* the coalescence variant takes the likelihood's step f[N/2] - f[N/2-1]; the others f[1] - f[0] */
    const double df = (mode == 2) ? f[N / 2] - f[N / 2 - 1] : f[1] - f[0];
    hipLaunchKernelGGL(gwh_maxll_reduce_kernel, dim3(1), dim3(GWH_MX_TPB), 0, st, N, 1, nblk, mode,
                       E.fft.as<const double2>(), E.mxpart.as<const double>(), 4.0 * df / 3.0, df,
                       E.mxres.as<double>());
    if (hipGetLastError() != hipSuccess) return nan_with(GWH_ERR_HIP, "reduce launch failed");
    double r[3] = {0, 0, 0};
    if (hipMemcpyAsync(r, E.mxres.p, 3 * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return nan_with(GWH_ERR_HIP, "readback failed");
    if (mode == 2) {
        if (tc) *tc = r[1];
        if (phiRef) *phiRef = r[2];
    }
    return r[0];
}

} // namespace

extern "C" {

double gwh_maximized_log_likelihood_aligned_internal(const double *data, const double *psd, const double *frequencies,
                                                     const double *response, size_t length)
{
    return given(data, psd, frequencies, response, nullptr, length, 0, nullptr, nullptr);
}

double gwh_maximized_log_likelihood_unaligned_internal(const double *data, const double *psd,
                                                       const double *frequencies, const double *hplus,
                                                       const double *hcross, size_t length)
{
    return given(data, psd, frequencies, hplus, hcross, length, 1, nullptr, nullptr);
}

double gwh_maximized_coal_log_likelihood_internal(const double *data, const double *psd, const double *frequencies,
                                                  const double *response, size_t length, double *tc, double *phiRef)
{
    return given(data, psd, frequencies, response, nullptr, length, 2, tc, phiRef);
}

int gwh_maximized_loglike_batch(int event, int model, const gwh_params *params, int n, double *out)
{
    Engine &E = gwh_engine();
    std::lock_guard<std::recursive_mutex> lk(E.mu);
    int rc = gwh_ensure_init();
    if (rc != GWH_OK) return rc;
    if (!params || !out || n < 0) return gwh_set_error(GWH_ERR_INVALID, "bad argument");
    if (n == 0) return GWH_OK;
    if (!gwh_model_valid(model)) return gwh_set_error(GWH_ERR_UNSUPPORTED, "bad model");
    if (event < 0 || event >= GWH_MAX_EVENTS - 1 || !E.events[event].live)
        return gwh_set_error(GWH_ERR_INVALID, "unknown event");
    const Event &ev = E.events[event];
    if (ev.method != GWH_SIMPSONS || ev.log10F)
        return gwh_set_error(GWH_ERR_UNSUPPORTED, "the maximised likelihoods need a Simpson-rule event");
    const int N = ev.N, D = ev.D;
    const bool pv2 = gwh_is_pv2(model);
    const int P = pv2 ? 2 : 1;
    const int nblk = (N + GWH_MX_TPB - 1) / GWH_MX_TPB;
    const size_t per_inst = (size_t)D * P * N * 16;
    int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)n, GWH_MX_BYTES / per_inst));
    if (chunk > 65535) chunk = 65535;
    /* the extrinsic values the intrinsic likelihood fixes (mcmc_gw.cpp:2611-2617 IMRPhenomD, :2660-2666 Pv2);
     * the statistics are invariant under the IMRPhenomD ones (amplitude, phase, time), not under Pv2's */
    std::vector<gwh_params> q(params, params + n);
    for (auto &p : q) {
        p.psi = 0;
        p.phiRef = 1;
        p.incl = 0;
        p.tc = 1;
        p.f_ref = pv2 ? 20. : 10.;
    }
    hipStream_t st = E.stream;
    if ((rc = E.params.ensure((size_t)chunk * sizeof(gwh_params))) != GWH_OK ||
        (rc = E.coef.ensure((size_t)chunk * sizeof(InstCoef))) != GWH_OK ||
        (rc = E.fft.ensure((size_t)chunk * per_inst)) != GWH_OK ||
        (rc = E.mxpart.ensure((size_t)chunk * D * nblk * 3 * 8)) != GWH_OK ||
        (rc = E.mxres.ensure((size_t)chunk * 8)) != GWH_OK)
        return rc;
    const double dfac = ev.df_like / 3.0;
    const double df = ev.df_fisher;                      /* f[1] - f[0] */
    const double hh_scale = ev.df_fisher / ev.df_like;  /* the stored Simpson step -> f[1] - f[0] */
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int m = std::min(chunk, n - i0);
        GWH_HIP(hipMemcpyAsync(E.params.p, q.data() + i0, (size_t)m * sizeof(gwh_params), hipMemcpyHostToDevice, st));
        if ((rc = gwh_setup_launch(nullptr, event, model, E.params.as<const gwh_params>(), m, E.coef.as<InstCoef>(),
                                   st)) != GWH_OK)
            return rc;
        dim3 g(nblk, m);
        if (pv2)
            hipLaunchKernelGGL(gwh_maxll_fill_kernel<GWH_IMRPHENOMPV2>, g, dim3(GWH_MX_TPB), 0, st,
                               (const EventDesc *)E.d_events, event, E.coef.as<const InstCoef>(), dfac,
                               E.fft.as<double2>(), E.mxpart.as<double>());
        else
            hipLaunchKernelGGL(gwh_maxll_fill_kernel<GWH_IMRPHENOMD>, g, dim3(GWH_MX_TPB), 0, st,
                               (const EventDesc *)E.d_events, event, E.coef.as<const InstCoef>(), dfac,
                               E.fft.as<double2>(), E.mxpart.as<double>());
        GWH_HIP(hipGetLastError());
        if ((rc = fft_forward(E.fft.as<double2>(), (size_t)N, (size_t)m * D * P, st)) != GWH_OK) return rc;
        hipLaunchKernelGGL(gwh_maxll_reduce_kernel, dim3(m), dim3(GWH_MX_TPB), 0, st, N, D, nblk, pv2 ? 1 : 0,
                           E.fft.as<const double2>(), E.mxpart.as<const double>(), hh_scale, df,
                           E.mxres.as<double>());
        GWH_HIP(hipGetLastError());
        GWH_HIP(hipMemcpyAsync(out + i0, E.mxres.p, (size_t)m * 8, hipMemcpyDeviceToHost, st));
    }
    GWH_HIP(hipStreamSynchronize(st));
    return GWH_OK;
}

} // extern "C"
