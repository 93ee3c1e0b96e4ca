/*
 * gwh_model.h -- per-instance setup and per-bin evaluation of IMRPhenomD and
 * IMRPhenomPv2 on the GPU (gfx950).
 *
 * The setup (FP64; one lane of each of four waves per instance) folds everything that does not
 * depend on frequency into one InstCoef record: the 19 calibration fits, the
 * ringdown/damping frequencies, the C1 connection coefficients of amplitude
 * and phase, the phase/time alignment, the precession constants, and for each
 * detector the complex projection constant and the time-of-flight phase slope.
 * The per-bin kernels then only evaluate the closed forms of one region.
 *
 * Reference: src/IMRPhenomD.cpp, src/IMRPhenomP.cpp, src/util.cpp,
 * src/waveform_generator.cpp, src/detector_util.cpp, src/waveform_util.cpp.
 */
#ifndef GWH_MODEL_H
#define GWH_MODEL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gwh_tables.h"
#include "gwh_trig.h"

#include "../../include/gwat_hip.h"

#define GWH_MAXDET 4
#define GWH_PI 3.141592653589793238462643383279502884
#define GWH_2PI 6.283185307179586476925286766559005768
#define GWH_MSOL_SEC 4.925491025543575903411922162094833998e-6
#define GWH_C_SI 299792458.
#define GWH_MPC_SI 3.085677581491367278913937957796471611e22
#define GWH_GAMMA_E 0.5772156649015328606065120900824024310421
#define GWH_LOG_64 4.15888308336 /* the reference's truncated ln 64 (IMRPhenomD.cpp:25) */
#define GWH_SQRT_6 2.44948974278317788

/* API model ids -> the kernel family (IMRPhenomD or IMRPhenomPv2) and the modification of its aligned-spin
 * phase */
__host__ __device__ inline bool gwh_model_valid(int m) { return m >= GWH_IMRPHENOMD && m <= GWH_GIMRPHENOMPV2; }
__host__ __device__ inline bool gwh_is_pv2(int m)
{
    return m == GWH_IMRPHENOMPV2 || m == GWH_PPE_IMRPHENOMPV2_INSPIRAL || m == GWH_PPE_IMRPHENOMPV2_IMR ||
           m == GWH_GIMRPHENOMPV2;
}
/* modification: 0 none, 1 ppE inspiral, 2 ppE all regions, 3 gIMR */
__host__ __device__ inline int gwh_ppe_variant(int m)
{
    if (m == GWH_PPE_IMRPHENOMD_INSPIRAL || m == GWH_PPE_IMRPHENOMPV2_INSPIRAL) return 1;
    if (m == GWH_PPE_IMRPHENOMD_IMR || m == GWH_PPE_IMRPHENOMPV2_IMR) return 2;
    if (m == GWH_GIMRPHENOMD || m == GWH_GIMRPHENOMPV2) return 3;
    return 0;
}

/* ---- theories mapped onto the ppE terms (ppE_utilities.cpp:158-360 assign_mapping, the beta functions
 * :408-735, applied in prep_source_parameters, waveform_generator.cpp:1379-1404) ---- */

/* redshift of a luminosity distance in Mpc (Z_from_DL, util.cpp:356-382): the segment whose upper bound exceeds
 * D_L holds a power series in sqrt(D_L) (cosmology_interpolation_function, util.cpp:502-512); -1 past the table */
__host__ __device__ inline double gwh_z_from_dl(double dl_mpc, int cosmo)
{
    for (int s = 0; s < GWH_COSMO_SEG; s++) {
        if (dl_mpc < gwh_cosmo_dl_bound[cosmo][s + 1]) {
            const double r = sqrt(dl_mpc);
            double z = gwh_cosmo_dz[cosmo][s][0], rk = 1.0;
            for (int k = 1; k < GWH_COSMO_DEG; k++) {
                rk *= r;
                z += gwh_cosmo_dz[cosmo][s][k] * rk;
            }
            return z;
        }
    }
    return -1.0;
}

/* ppE terms a theory maps to (the mapping's Nmod); each reads as many theory parameters from ppe_beta */
__host__ __device__ inline int gwh_theory_nterms(int theory, int nmod)
{
    if (theory >= GWH_THEORY_EDGB_GHOV1 && theory <= GWH_THEORY_EDGB_GHOV3) return 2;
    if (theory == GWH_THEORY_PNSERIES || theory == GWH_THEORY_PPEALT) return nmod;
    return 1;
}

/* the scalar charges of dCS (dCS_phase_factor, ppE_utilities.cpp:487-523) and EdGB (EdGB_phase_factor,
 * :589-616) of a black hole of dimensionless spin chi; 0 below the reference's comparison threshold */
__host__ __device__ inline double gwh_dcs_charge(double chi)
{
    if (fabs(chi) < 1e-10) return 0.0;
    const double c2 = chi * chi, r = sqrt(1. - c2);
    return (2. + 2. * c2 * c2 - 2. * r - c2 * (3. - 2. * r)) / (2. * c2 * chi);
}
__host__ __device__ inline double gwh_edgb_charge(double chi)
{
    if (fabs(chi) < 1e-10) return 0.0;
    const double c2 = chi * chi;
    return 2. * (sqrt(1. - c2) - 1. + c2) / c2;
}

/* The ppE exponents b[] and coefficients beta[] of a theory, from the source in seconds (total mass M, chirp mass
 * Mc, eta as the waveform uses them) and the theory parameters ppe_beta[] (ppe_b[] for PNSeries / ppEAlt).
 * Returns the term count, or -1 for an unknown theory / cosmology or too few theory parameters. */
__host__ __device__ inline int gwh_theory_terms(const gwh_params &p, double M, double Mc, double eta, double *b,
                                                double *beta)
{
    const int th = p.ppe_theory, nm = p.ppe_nmod;
    if (th <= GWH_THEORY_NONE || th >= GWH_THEORY_COUNT || p.cosmology < 0 || p.cosmology >= GWH_NCOSMO) return -1;
    const int n = gwh_theory_nterms(th, nm);
    if (n < 1 || n > GWH_PPE_MAX || nm < n) return -1;
    const double a0 = p.ppe_beta[0];
    if (th == GWH_THEORY_PNSERIES || th == GWH_THEORY_PPEALT) { /* PNSeries_beta / ppEAlt_beta (:408-443) */
        const double m1 = p.mass1 * GWH_MSOL_SEC, m2 = p.mass2 * GWH_MSOL_SEC;
        const double ratio = (m1 + m2) / (pow(m1 * m2, 3. / 5.) / pow(m1 + m2, 1. / 5.));
#pragma unroll
        for (int i = 0; i < GWH_PPE_MAX; i++) { /* constant indices: the terms stay in registers */
            if (i >= n) break;
            b[i] = p.ppe_b[i];
            const double amp = (th == GWH_THEORY_PNSERIES && i > 0) ? a0 * p.ppe_beta[i] : p.ppe_beta[i];
            beta[i] = amp * pow(ratio, b[i] / 3.);
        }
        return n;
    }
    const double z = gwh_z_from_dl(p.DL, p.cosmology);
    const double Ms = M / (1. + z), Ms4 = Ms * Ms * Ms * Ms;
    const double chi_s = 0.5 * (p.spin1[2] + p.spin2[2]), chi_a = 0.5 * (p.spin1[2] - p.spin2[2]);
    const double chi1 = chi_s + chi_a, chi2 = chi_s - chi_a;
    const double ep = pow(eta, 3. / 5.), sq = sqrt(1. - 4. * eta);
    const double m1 = 0.5 * (Mc / ep + sq * Mc / ep), m2 = 0.5 * (Mc / ep - sq * Mc / ep); /* calculate_mass1/2 */
    /* EdGB_phase_factor (:589-616) */
    const double e1 = gwh_edgb_charge(chi1), e2 = gwh_edgb_charge(chi2), ed = m1 * m1 * e2 - m2 * m2 * e1;
    const double edgb_f = (-5. / 7168.) * (ed * ed) / (M * M * M * M * pow(eta, 18. / 5.));
    const double scale = 16. * GWH_PI * a0 / Ms4;
    /* (3 - 26 eta + 34 eta^2) / (eta^(2/5) (1 - 2 eta)) of the mass-loss terms (:667-695) */
    const double mdot_f = (25. / 851968.) * ((3. - 26. * eta + 34. * eta * eta) / (pow(eta, 2. / 5.) * (1 - 2 * eta)));
    switch (th) {
    case GWH_THEORY_DCS: { /* dCS_beta, dCS_phase_factor (:473-523) */
        const double d1 = gwh_dcs_charge(chi1), d2 = gwh_dcs_charge(chi2), m = m1 + m2;
        const double e145 = pow(eta, 14. / 5.), dd = m1 * d2 - m2 * d1;
        double g = 0;
        g += (-5. / 8192.) / e145 * pow(dd, 2.) / (m * m);
        g += (15075. / 114688.) / e145 *
             (m2 * m2 * chi1 * chi1 - 350. / 201. * m1 * m2 * chi1 * chi2 + m1 * m1 * chi2 * chi2) / (m * m);
        b[0] = -1;
        beta[0] = scale * g;
        break;
    }
    case GWH_THEORY_EDGB: /* EdGB_beta (:573-582) */
        b[0] = -7;
        beta[0] = scale * edgb_f;
        break;
    case GWH_THEORY_EDGB_HO_LO: /* EdGB_HO_0PN_beta (:530-542) */
        b[0] = -7;
        beta[0] = -5. * (16. * GWH_PI * a0) / Ms4 / 7168. / pow(eta, 18. / 5.) * (4 * eta - 1);
        break;
    case GWH_THEORY_EDGB_GHOV1:
    case GWH_THEORY_EDGB_GHOV2:
    case GWH_THEORY_EDGB_GHOV3: { /* EdGB_beta and EdGB_GHO_betav1..3 (:622-664) */
        b[0] = -7;
        b[1] = -5;
        beta[0] = scale * edgb_f;
        double g = p.ppe_beta[1];
        if (th != GWH_THEORY_EDGB_GHOV1) g = edgb_f * g;
        if (th == GWH_THEORY_EDGB_GHOV3) g = g * (3715. / 756. + 55. * eta / 9.);
        beta[1] = scale * g;
        break;
    }
    case GWH_THEORY_EXTRADIMENSION: { /* ExtraDimension_beta (:667-679): m-dot of each hole in the brane world */
        const double l10 = 10.e-6 / GWH_C_SI, year = 31557600.;
        const double q1 = GWH_MSOL_SEC * (1 + z) / (p.mass1 * GWH_MSOL_SEC), q2 = GWH_MSOL_SEC * (1 + z) / (p.mass2 * GWH_MSOL_SEC);
        const double md1 = -2.8e-7 * (q1 * q1) * a0 / (l10 * l10) * GWH_MSOL_SEC / year;
        const double md2 = -2.8e-7 * (q2 * q2) * a0 / (l10 * l10) * GWH_MSOL_SEC / year;
        b[0] = -13;
        beta[0] = (md1 + md2) * mdot_f;
        break;
    }
    case GWH_THEORY_BHEVAPORATION: /* BHEvaporation_beta (:685-692) */
        b[0] = -13;
        beta[0] = a0 * mdot_f;
        break;
    case GWH_THEORY_TVG: /* TVG_beta (:698-704) */
        b[0] = -13;
        beta[0] = (-25. / 65526.) * (a0 * Mc / (1 + z));
        break;
    case GWH_THEORY_DIPRAD: /* DipRad_beta (:710-716) */
        b[0] = -7;
        beta[0] = (-3. / 224.) * pow(eta, 2. / 5.) * a0;
        break;
    case GWH_THEORY_MODDISPERSION: { /* ModDispersion_beta (:757-772) with DL_from_Z_MD (:789-812) */
        const double al = (p.ppe_b[0] + 3.) / 3.;
        int ia = -1;
        for (int a = 0; a < GWH_MD_NALPHA; a++)
            if (ia < 0 && fabs(al - gwh_md_alpha[a]) < 1e-10) ia = a;
        if (ia < 0) return -1; /* a dispersion power without a distance fit */
        double D = -1;
        for (int s = 0; s < GWH_MD_SEG; s++) {
            if (z < gwh_md_z_bound[ia][s + 1]) {
                D = 0;
                for (int k = 0; k < GWH_MD_DEG; k++) D += gwh_md_dz[ia][s][k] * pow(z, -3.5 + k * 0.5);
                break;
            }
        }
        D *= GWH_MPC_SI / GWH_C_SI;
        b[0] = p.ppe_b[0];
        beta[0] = (pow(GWH_PI, 2. - al) / (1. - al)) * (D * a0 / pow(4.135667696e-15, 2. - al)) *
                  (pow(Mc, 1. - al) / pow(1 + z, 1. - al));
        break;
    }
    default: /* GWH_THEORY_NONCOMM, NonComm_beta (:721-726) */
        b[0] = -1;
        beta[0] = (-75. / 256.) * pow(eta, -4. / 5.) * (2. * eta - 1.) * a0;
        break;
    }
    return n;
}

/* Frequency-independent constants of one instance (one parameter vector on
 * one event).  All dimensionful quantities in seconds / Hz. */
struct InstCoef {
    double M, eta, fcut;
    double f1, f3, f1p, f2p; /* amplitude and phase transition frequencies */
    double fRD, fdamp;
    double amp_pref;         /* A0 M^{7/6} (Pv2: / (2 sqrt(5/64 pi)) / 2) */
    double ca[7];            /* inspiral amplitude PN coefficients times pi^{i/3}: polynomial in u = (Mf)^{1/3} */
    double rho[3];           /* rho_1..3 (powers u^7, u^8, u^9) */
    double delta[5];         /* intermediate amplitude polynomial in Mf */
    double mr_k;             /* gamma1 gamma3 fdamp / M */
    double mr_g2;            /* gamma2 / (gamma3 fdamp) */
    double mr_w2;            /* (gamma3 fdamp)^2 */
    double K;                /* 3 / (128 eta) */
    double cp[8];            /* TF2 phase coefficients in v = (pi M f)^{1/3}; cp[5], cp[6] are log dependent */
    double c8, c9, c10, c11; /* cp5 = c8 + c9 ln v^3, cp6 = c10 - c11 ln v^3 */
    double sig[4];           /* sigma_1/eta, (3/4) sigma_2/eta, (3/5) sigma_3/eta, (1/2) sigma_4/eta */
    double beta[4];          /* beta_k / eta (beta_3 / (3 eta)) */
    double alpha[6];         /* alpha_k / eta ((4/3) alpha_3 / eta) ; alpha[5] = alpha_5 fRD */
    double tc_total, phic, f_ref; /* phi = Psi - tc_total (f - f_ref) - phic (Pv2: + t_corr term) */
    double t_corr_2pi;       /* Pv2: 2 pi t_corr */
    /* Pv2 precession */
    double SP, SL, L_c1, L_c2; /* L2PN: eta (1 + L_c1 x + L_c2 x^2) / sqrt(x) */
    double a_c[5], e_c[5];   /* Euler alpha / epsilon coefficients */
    double alpha_add;        /* alpha0 - alpha(f_ref) */
    double eps_add;          /* -epsilon(f_ref) */
    double ylm[5];           /* -2Y_2m(thetaJN, 0), m = -2..2 */
    /* per detector: PhenomD: (Re K, Im K, |K|^2, 2 pi dt) ; Pv2: (P, Q, 0, 2 pi dt) */
    double det[GWH_MAXDET][4];
    double wf_a, wf_b; /* polarisation constants: PhenomD ((1+cos^2 i)/2, cos i); Pv2 (cos 2 zeta, sin 2 zeta) */
    /* powers of the total mass that pair with the tabulated powers of each bin: (Mf)^p = M^p f^p */
    double M13, Mm13, Mm76, M34, lgM, lgPiM;
    double inv_fdamp; /* 1 / fdamp */
    /* ppE phase terms: sum_i ppe_c[i] f^ppe_e[i], ppe_c = beta (pi Mc)^(b/3), ppe_e = b/3; added to the
     * inspiral phase (ppe_imr = 0) or to every region (ppe_imr = 1) */
    int ppe_n, ppe_imr;
    double ppe_e[GWH_PPE_MAX], ppe_c[GWH_PPE_MAX];
    /* gIMR: inspiral phase terms K (cp[1] v^-4 + sum_i gimr_neg[i + 4] v^(i - 5), i = -4..-1), v = (pi M f)^(1/3) */
    int gimr;
    double gimr_neg[4];
};

/* Powers of one frequency bin that every instance needs (log f, f^(1/3), f^(-1/3), f^(-7/6), f^(3/4)).
 * Resident events tabulate them once; single calls evaluate them per bin with the same function. */
struct BinPow {
    double lf, f13, fm13, fm76, f34;
};

__device__ __forceinline__ BinPow gwh_bin_powers(double f)
{
    BinPow b;
    b.lf = log(f);
    b.f13 = cbrt(f);
    b.fm13 = 1.0 / b.f13;
    b.fm76 = 1.0 / (f * sqrt(b.f13));
    b.f34 = sqrt(sqrt(f * f * f));
    return b;
}

/* QNM spline table in device memory: knots, values and S''/2 of the natural spline */
struct QnmTable {
    const double *spin, *fring, *fdamp, *c_ring, *c_damp;
    int n;
    double x0, x1; /* first and last spin knot */
};

/* detector geometry in device memory */
struct DetGeom {
    double tensor[9];
    double loc[3];
    double gfac;
};

/* ------------------------------------------------------------------------- */
/* FP64 setup helpers (device)                                                 */
/* ------------------------------------------------------------------------- */

__device__ __forceinline__ double gwh_spline_eval(const double *x, const double *y, const double *c, int n,
                                                  double xq)
{
    if (!(xq >= x[0] && xq <= x[n - 1])) return __longlong_as_double(0x7ff8000000000000LL);
    /* interval with x[lo] <= xq < x[lo + 1] (lo = n - 2 at the last knot), the one a bisection finds:
     * start from the uniform-spacing guess and walk (a step or two on the near-uniform spin grid) */
    int lo = (int)((xq - x[0]) / (x[n - 1] - x[0]) * (n - 1));
    lo = lo < 0 ? 0 : (lo > n - 2 ? n - 2 : lo);
    while (lo > 0 && x[lo] > xq) lo--;
    while (lo < n - 2 && x[lo + 1] <= xq) lo++;
    double h = x[lo + 1] - x[lo];
    double b = (y[lo + 1] - y[lo]) / h - h * (c[lo + 1] + 2.0 * c[lo]) / 3.0;
    double d = (c[lo + 1] - c[lo]) / (3.0 * h);
    double dx = xq - x[lo];
    return y[lo] + dx * (b + dx * (c[lo] + dx * d));
}

/* both QNM splines (M f_ring and M f_damp) at spin a, with the interval x[lo] <= a < x[lo + 1] (lo = n - 2 at
 * the last knot) that a bisection finds.  On the table's near-uniform spin grid that interval lies within one
 * knot of the uniform-spacing guess, so the knots and coefficients of the four-knot window around the guess
 * are loaded together (one memory round trip) and the interval is picked among them; a window that misses it
 * falls back to the walk of gwh_spline_eval.  Same arithmetic as gwh_spline_eval. */
__device__ __forceinline__ void gwh_qnm_eval(const QnmTable &q, double a, double &fring, double &fdamp)
{
    const int n = q.n;
    if (!(a >= q.x0 && a <= q.x1)) {
        fring = fdamp = __longlong_as_double(0x7ff8000000000000LL);
        return;
    }
    int g = (int)((a - q.x0) / (q.x1 - q.x0) * (n - 1));
    g = g < 0 ? 0 : (g > n - 2 ? n - 2 : g);
    const int b = (g - 1 < 0) ? 0 : (g - 1 > n - 4 ? n - 4 : g - 1);
    double xs[4], yr[4], cr[4], yd[4], cd[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        xs[k] = q.spin[b + k];
        yr[k] = q.fring[b + k];
        cr[k] = q.c_ring[b + k];
        yd[k] = q.fdamp[b + k];
        cd[k] = q.c_damp[b + k];
    }
    if (!(xs[0] <= a && (a < xs[3] || b == n - 4))) {
        fring = gwh_spline_eval(q.spin, q.fring, q.c_ring, n, a);
        fdamp = gwh_spline_eval(q.spin, q.fdamp, q.c_damp, n, a);
        return;
    }
    double x0 = xs[0], x1 = xs[1], r0 = yr[0], r1 = yr[1], rc0 = cr[0], rc1 = cr[1], d0 = yd[0], d1 = yd[1],
           dc0 = cd[0], dc1 = cd[1];
#pragma unroll
    for (int k = 1; k < 3; k++)
        if (xs[k] <= a) {
            x0 = xs[k];
            x1 = xs[k + 1];
            r0 = yr[k];
            r1 = yr[k + 1];
            rc0 = cr[k];
            rc1 = cr[k + 1];
            d0 = yd[k];
            d1 = yd[k + 1];
            dc0 = cd[k];
            dc1 = cd[k + 1];
        }
    const double h = x1 - x0, dx = a - x0;
    double bb = (r1 - r0) / h - h * (rc1 + 2.0 * rc0) / 3.0;
    double dd = (rc1 - rc0) / (3.0 * h);
    fring = r0 + dx * (bb + dx * (rc0 + dx * dd));
    bb = (d1 - d0) / h - h * (dc1 + 2.0 * dc0) / 3.0;
    dd = (dc1 - dc0) / (3.0 * h);
    fdamp = d0 + dx * (bb + dx * (dc0 + dx * dd));
}

/* 19 calibration fits (IMRPhenomD.cpp:852-870) */
__device__ __forceinline__ double gwh_fit(const double *tab, int row, double eta, double x)
{
    const double *p = tab + 11 * row;
    return p[0] + p[1] * eta + x * (p[2] + p[3] * eta + p[4] * eta * eta) +
           x * x * (p[5] + p[6] * eta + p[7] * eta * eta) + x * x * x * (p[8] + p[9] * eta + p[10] * eta * eta);
}

__device__ __forceinline__ double gwh_final_spin0815(double eta, double chi1, double chi2)
{
    double Seta = sqrt(1.0 - 4.0 * eta);
    double m1 = 0.5 * (1.0 + Seta), m2 = 0.5 * (1.0 - Seta);
    double s = m1 * m1 * chi1 + m2 * m2 * chi2;
    double eta2 = eta * eta, eta3 = eta2 * eta, s2 = s * s, s3 = s2 * s;
    return eta * (3.4641016151377544 - 4.399247300629289 * eta + 9.397292189321194 * eta2 -
                  13.180949901606242 * eta3 +
                  s * ((1.0 / eta - 0.0850917821418767 - 5.837029316602263 * eta) +
                       (0.1014665242971878 - 2.0967746996832157 * eta) * s +
                       (-1.3546806617824356 + 4.108962025369336 * eta) * s2 +
                       (-0.8676969352555539 + 2.064046835273906 * eta) * s3));
}

__device__ __forceinline__ double gwh_erad0815(double eta, double chi1, double chi2)
{
    double Seta = sqrt(1.0 - 4.0 * eta);
    double m1 = 0.5 * (1.0 + Seta), m2 = 0.5 * (1.0 - Seta);
    double m1s = m1 * m1, m2s = m2 * m2;
    double s = (m1s * chi1 + m2s * chi2) / (m1s + m2s);
    double eta2 = eta * eta, eta3 = eta2 * eta;
    return (eta * (0.055974469826360077 + 0.5809510763115132 * eta - 0.9606726679372312 * eta2 +
                   3.352411249771192 * eta3) *
            (1. + (-0.0030302335878845507 - 2.0066110851351073 * eta + 7.7050567802399215 * eta2) * s)) /
           (1. + (-0.6714403054720589 - 1.4756929437702908 * eta + 7.304676214885011 * eta2) * s);
}

/* Everything the setup needs while building the record */
struct SetupState {
    double m1, m2, M, eta, chi1, chi2, chi_s, chi_a, delta, chi_pn, chi_eff, Mc, A0;
    double fRD, fdamp, f1, f3, f1p, f2p;
    double rho[3], v2, g[3], sigma[5], beta[4], alpha[6];
    double pa[7];  /* amplitude PN coefficients (reference numbering) */
    double pp[12]; /* phase PN coefficients (reference numbering), [5],[6] dynamic */
    double delta_c[5];
    int ppe_n, ppe_imr;
    double ppe_e[GWH_PPE_MAX], ppe_c[GWH_PPE_MAX];
    int gimr;
    double gimr_neg[4]; /* delta_phi of PN orders -4 .. -1 */
};

/* 2.5PN phase coefficient of IMRPhenomD (IMRPhenomD.cpp:1016-1018; also the coefficient of its log term) */
__device__ __forceinline__ double gwh_pn_c8_gr(double eta, double dm, double ca, double cs)
{
    const double eta2 = eta * eta;
    return (38645. * GWH_PI / 756 - 65. * GWH_PI * eta / 9 + dm * (-732985. / 2268 - 140. * eta / 9) * ca +
            (-732985. / 2268 + 24260. * eta / 81 + 340. * eta2 / 9) * cs);
}

/* gIMR inspiral phase terms beyond the IMRPhenomD polynomial (gIMRPhenomD.cpp:148-184): the order-1 coefficient
 * (zero in IMRPhenomD) and 3/(128 eta) delta_i (pi M f)^((i - 5)/3) for i = -4..-1, with their f derivative */
__device__ __forceinline__ double gwh_setup_gimr(const SetupState &S, double f, double *deriv)
{
    const double K = 3. / (128. * S.eta);
    const double x = GWH_PI * S.M * f, iv = 1.0 / cbrt(x);
    double iv2 = iv * iv, iv4 = iv2 * iv2;
    double t = S.pp[1] * iv4, d = S.pp[1] * (-4. / 3.) * iv4;
    double p = iv4 * iv; /* iv^5 */
    for (int i = -1; i >= -4; i--) {
        p *= iv; /* iv^(5 - i) */
        t += S.gimr_neg[i + 4] * p;
        d += S.gimr_neg[i + 4] * ((i - 5.) / 3.) * p;
    }
    if (deriv) *deriv = K * d / f;
    return K * t;
}

/* ppE phase terms at f and their frequency derivative (ppE_IMRPhenomD.cpp:22-48): beta (pi Mc f)^(b/3) and
 * (b/3) beta (pi Mc f)^(b/3) / f */
__device__ __forceinline__ double gwh_setup_ppe(const SetupState &S, double f, double *deriv)
{
    double t = 0, d = 0;
    const double lf = log(f);
    /* a fixed trip count with a guard keeps every index a compile-time constant (registers, no stack) */
#pragma unroll
    for (int i = 0; i < GWH_PPE_MAX; i++) {
        if (i >= S.ppe_n) break;
        const double v = S.ppe_c[i] * exp(S.ppe_e[i] * lf);
        t += v;
        d += S.ppe_e[i] * v;
    }
    if (deriv) *deriv = d / f;
    return t;
}

/* TF2 + sigma inspiral phase at f (IMRPhenomD.cpp:1364-1399) */
__device__ __forceinline__ double gwh_setup_phase_ins(const SetupState &S, double f)
{
    double x = GWH_PI * S.M * f;
    double v = cbrt(x);
    double lx = log(x);
    double c5 = S.pp[8] + lx * S.pp[9];
    double c6 = S.pp[10] - S.pp[11] * lx;
    double v2 = v * v;
    double poly = S.pp[0] + v2 * (S.pp[2] + v * (S.pp[3] + v * (S.pp[4] + v * (c5 + v * (c6 + v * S.pp[7])))));
    double Mf = S.M * f;
    double u = cbrt(Mf);
    double Mf43 = Mf * u, Mf53 = Mf43 * u;
    double ph = -GWH_PI / 4. + 3. / (128. * S.eta) * poly / (v2 * v2 * v) +
                (1. / S.eta) * (S.sigma[1] * Mf + 0.75 * S.sigma[2] * Mf43 + 0.6 * S.sigma[3] * Mf53 +
                                0.5 * S.sigma[4] * Mf * Mf);
    if (S.ppe_n > 0) ph += gwh_setup_ppe(S, f, nullptr);
    if (S.gimr) ph += gwh_setup_gimr(S, f, nullptr);
    return ph;
}

/* d/df of the inspiral phase (IMRPhenomD.cpp:1426-1460) */
__device__ __forceinline__ double gwh_setup_dphase_ins(const SetupState &S, double f)
{
    double M = S.M;
    double x = GWH_PI * M * f;
    double lx = log(x);
    double c[8];
    for (int i = 0; i < 8; i++) c[i] = S.pp[i];
    c[5] = S.pp[8] + lx * S.pp[9];
    c[6] = S.pp[10] - S.pp[11] * lx;
    double K = 3. / (128. * S.eta);
    /* x^((i - 8) / 3) as powers of x^(-1/3) */
    double v = cbrt(x), iv = 1.0 / v;
    double sum = 0, pw = iv;
#pragma unroll
    for (int i = 7; i >= 0; i--) {
        sum += c[i] * ((i - 5.) / 3.) * pw;
        pw *= iv;
    }
    double d = K * (GWH_PI * M) * sum + K * (S.pp[9] - S.pp[11] * v) / f;
    /* M^(4/3) f^(1/3) = M (Mf)^(1/3), M^(5/3) f^(2/3) = M (Mf)^(2/3) */
    double u = cbrt(M * f);
    double sg = S.sigma[1] * M + S.sigma[2] * M * u + S.sigma[3] * M * u * u + S.sigma[4] * M * M * f;
    double r = d + sg / S.eta;
    if (S.ppe_n > 0) {
        double dt;
        gwh_setup_ppe(S, f, &dt);
        r += dt;
    }
    if (S.gimr) { /* the order-1 coefficient is already in the polynomial derivative above */
        double dt;
        gwh_setup_gimr(S, f, &dt);
        const double iv4 = iv * iv * iv * iv;
        r += dt - K * S.pp[1] * (-4. / 3.) * iv4 / f;
        /* The reference differentiates the log terms of c5 and c6 with their GR coefficients whatever gIMR did
         * to them (assign_nonstatic_pn_phase_coeff_deriv, IMRPhenomD.cpp:1068-1084, reached from gIMRPhenomD's
         * Dphase_ins through the GR class, gIMRPhenomD.cpp:167-184): with phi orders 8 and 9 set, the
         * connection slope at f1 is not the inspiral phase's own slope. */
        const double c9gr = gwh_pn_c8_gr(S.eta, S.delta, S.chi_a, S.chi_s), c11gr = 6848. / 63.;
        r += K * ((c9gr - S.pp[9]) - (c11gr - S.pp[11]) * v) / f;
    }
    return r;
}

/* the ppE_IMRPhenomD_IMR terms of the intermediate and merger-ringdown phases (ppE_IMRPhenomD.cpp:54-103) */
__device__ __forceinline__ double gwh_setup_ppe_imr(const SetupState &S, double f, bool deriv)
{
    if (S.ppe_n <= 0 || !S.ppe_imr) return 0;
    double d;
    double t = gwh_setup_ppe(S, f, &d);
    return deriv ? d : t;
}

/* intermediate phase at f, given lMf = ln(M f) */
__device__ __forceinline__ double gwh_setup_phase_int_l(const SetupState &S, double f, double lMf)
{
    double Mf = S.M * f;
    return (1. / S.eta) * (S.beta[0] + S.beta[1] * Mf + S.beta[2] * lMf - S.beta[3] / 3. / (Mf * Mf * Mf)) +
           gwh_setup_ppe_imr(S, f, false);
}

__device__ __forceinline__ double gwh_setup_phase_int(const SetupState &S, double f)
{
    return gwh_setup_phase_int_l(S, f, log(S.M * f));
}

__device__ __forceinline__ double gwh_setup_dphase_int(const SetupState &S, double f)
{
    double M = S.M;
    return (1. / S.eta) * (S.beta[1] * M + S.beta[2] / f + S.beta[3] / (M * M * M * f * f * f * f)) +
           gwh_setup_ppe_imr(S, f, true);
}

__device__ __forceinline__ double gwh_setup_phase_mr(const SetupState &S, double f)
{
    double Mf = S.M * f;
    return (1. / S.eta) * (S.alpha[0] + S.alpha[1] * Mf - S.alpha[2] / Mf + (4. / 3.) * S.alpha[3] * sqrt(sqrt(Mf * Mf * Mf)) +
                           S.alpha[4] * atan((f - S.alpha[5] * S.fRD) / S.fdamp)) +
           gwh_setup_ppe_imr(S, f, false);
}

__device__ __forceinline__ double gwh_setup_dphase_mr(const SetupState &S, double f)
{
    double M = S.M;
    double t = (f - S.alpha[5] * S.fRD) / S.fdamp;
    return (S.alpha[4] / (S.fdamp * (1. + t * t)) + S.alpha[2] / (f * f * M) + M * (S.alpha[1] + S.alpha[3] / sqrt(sqrt(f * M)))) /
               S.eta +
           gwh_setup_ppe_imr(S, f, true);
}

__device__ __forceinline__ double gwh_setup_build_phase(const SetupState &S, double f)
{
    if (f < S.f1p) return gwh_setup_phase_ins(S, f);
    if (f > S.f2p) return gwh_setup_phase_mr(S, f);
    return gwh_setup_phase_int(S, f);
}

/* reduced inspiral amplitude and its frequency derivative (IMRPhenomD.cpp:1335-1420) */
__device__ __forceinline__ void gwh_setup_amp_ins(const SetupState &S, double f, double *val, double *der)
{
    double M = S.M;
    double x = GWH_PI * M * f, y = M * f;
    double vx = cbrt(x), vy = cbrt(y);
    double v = 0, d = 0, p = 1.0;
    for (int i = 0; i < 7; i++) {
        v += S.pa[i] * p;
        if (i > 0) d += S.pa[i] * (i / 3.) * p / x * (GWH_PI * M);
        p *= vx;
    }
    double y73 = y * y * vy, y83 = y73 * vy, y3 = y * y * y;
    v += S.rho[0] * y73 + S.rho[1] * y83 + S.rho[2] * y3;
    d += (S.rho[0] * (7. / 3.) * y73 + S.rho[1] * (8. / 3.) * y83 + S.rho[2] * 3. * y3) / y * M;
    *val = v;
    *der = d;
}

/* reduced merger-ringdown amplitude and its derivative (IMRPhenomD.cpp:1467-1532) */
__device__ __forceinline__ void gwh_setup_amp_mr(const SetupState &S, double f, double *val, double *der)
{
    double dl = f - S.fRD, w = S.g[2] * S.fdamp;
    double L = dl * dl + w * w;
    double a = S.g[0] * S.g[2] * S.fdamp * S.M * exp(-S.g[1] * dl / w) / (S.M * S.M * L);
    *val = a;
    *der = a * (-S.g[1] / w - 2.0 * dl / L);
}

/* ---- Pv2 precession helpers (IMRPhenomP.cpp) ---- */
__device__ __forceinline__ double gwh_l2pn(double eta, double x)
{
    double eta2 = eta * eta;
    return (eta * (1.0 + (1.5 + eta / 6.0) * x + (3.375 - (19.0 * eta) / 8. - eta2 / 24.0) * x * x)) / sqrt(x);
}

/* NNLO Euler-angle coefficients (IMRPhenomP.cpp:64-144) */
__device__ inline void gwh_euler_coeffs(double q, double chil, double chip, double *a, double *e)
{
    double m2 = q / (1. + q), m1 = 1. / (1. + q);
    double dm = m1 - m2;
    double eta = m1 * m2, eta2 = eta * eta, eta3 = eta2 * eta, eta4 = eta3 * eta;
    double chil2 = chil * chil, chip2 = chip * chip, chip4 = chip2 * chip2;
    double dm2 = dm * dm, dm3 = dm2 * dm;
    double m2_2 = m2 * m2, m2_3 = m2_2 * m2, m2_4 = m2_3 * m2, m2_5 = m2_4 * m2, m2_6 = m2_5 * m2, m2_7 = m2_6 * m2,
           m2_8 = m2_7 * m2;
    a[0] = (-0.18229166666666666 - (5 * dm) / (64. * m2));
    a[1] = ((-15 * dm * m2 * chil) / (128. * eta) - (35 * m2_2 * chil) / (128. * eta));
    a[2] = (-1.7952473958333333 - (4555 * dm) / (7168. * m2) - (15 * chip2 * dm * m2_3) / (128. * eta2) -
            (35 * chip2 * m2_4) / (128. * eta2) - (515 * eta) / 384. - (15 * dm2 * eta) / (256. * m2_2) -
            (175 * dm * eta) / (256. * m2));
    a[3] = -(35 * GWH_PI) / 48. - (5 * dm * GWH_PI) / (16. * m2) + (5 * dm2 * chil) / 16. + (5 * dm * m2 * chil) / 3. +
           (2545 * m2_2 * chil) / 1152. - (5 * chip2 * dm * m2_5 * chil) / (128. * eta3) -
           (35 * chip2 * m2_6 * chil) / (384. * eta3) + (2035 * dm * m2 * chil) / (21504. * eta) +
           (2995 * m2_2 * chil) / (9216. * eta);
    a[4] = (4.318908476114694 + (27895885 * dm) / (2.1676032e7 * m2) - (15 * chip4 * dm * m2_7) / (512. * eta4) -
            (35 * chip4 * m2_8) / (512. * eta4) - (485 * chip2 * dm * m2_3) / (14336. * eta2) +
            (475 * chip2 * m2_4) / (6144. * eta2) + (15 * chip2 * dm2 * m2_2) / (256. * eta) +
            (145 * chip2 * dm * m2_3) / (512. * eta) + (575 * chip2 * m2_4) / (1536. * eta) + (39695 * eta) / 86016. +
            (1615 * dm2 * eta) / (28672. * m2_2) - (265 * dm * eta) / (14336. * m2) + (955 * eta2) / 576. +
            (15 * dm3 * eta2) / (1024. * m2_3) + (35 * dm2 * eta2) / (256. * m2_2) + (2725 * dm * eta2) / (3072. * m2) -
            (15 * dm * m2 * GWH_PI * chil) / (16. * eta) - (35 * m2_2 * GWH_PI * chil) / (16. * eta) +
            (15 * chip2 * dm * m2_7 * chil2) / (128. * eta4) + (35 * chip2 * m2_8 * chil2) / (128. * eta4) +
            (375 * dm2 * m2_2 * chil2) / (256. * eta) + (1815 * dm * m2_3 * chil2) / (256. * eta) +
            (1645 * m2_4 * chil2) / (192. * eta));
    e[0] = a[0];
    e[1] = a[1];
    e[2] = (-1.7952473958333333 - (4555 * dm) / (7168. * m2) - (515 * eta) / 384. - (15 * dm2 * eta) / (256. * m2_2) -
            (175 * dm * eta) / (256. * m2));
    e[3] = -(35 * GWH_PI) / 48. - (5 * dm * GWH_PI) / (16. * m2) + (5 * dm2 * chil) / 16. + (5 * dm * m2 * chil) / 3. +
           (2545 * m2_2 * chil) / 1152. + (2035 * dm * m2 * chil) / (21504. * eta) + (2995 * m2_2 * chil) / (9216. * eta);
    e[4] = (4.318908476114694 + (27895885 * dm) / (2.1676032e7 * m2) + (39695 * eta) / 86016. +
            (1615 * dm2 * eta) / (28672. * m2_2) - (265 * dm * eta) / (14336. * m2) + (955 * eta2) / 576. +
            (15 * dm3 * eta2) / (1024. * m2_3) + (35 * dm2 * eta2) / (256. * m2_2) + (2725 * dm * eta2) / (3072. * m2) -
            (15 * dm * m2 * GWH_PI * chil) / (16. * eta) - (35 * m2_2 * GWH_PI * chil) / (16. * eta) +
            (375 * dm2 * m2_2 * chil2) / (256. * eta) + (1815 * dm * m2_3 * chil2) / (256. * eta) +
            (1645 * m2_4 * chil2) / (192. * eta));
}

__device__ __forceinline__ double gwh_euler_eval(const double *c, double v, double logw)
{
    /* the negative powers of v from one Newton-refined reciprocal (the two angles of a bin share it once inlined)
       instead of three FP64 division sequences; the sum as explicit fused multiply-adds, so every kernel that
       inlines it (the GR and the modified-gravity setup among them) rounds it the same way */
    const double iv = rcp_newton(v), iv2 = iv * iv;
    return fma(c[0], iv2 * iv, fma(c[1], iv2, fma(c[2], iv, fma(c[3], logw, c[4] * v))));
}

/* ------------------------------------------------------------------------- */
/* the setup proper, in parts that run on four waves                         */
/* ------------------------------------------------------------------------- */
/*
 * The frequency-independent constants of one instance are a long chain of FP64 work, and a wave issues it
 * at the same speed whatever the number of its active lanes.  The setup kernel therefore gives every
 * instance one lane in each of four waves (roles) that compute independent parts and exchange them through
 * one LDS record per instance (SetupShared):
 *
 *   phase 1  role 0  calibration fits, remnant spin, ringdown / damping and transition frequencies, the
 *                    merger-ringdown amplitude and its derivative at f3
 *            role 1  PN coefficients, the inspiral amplitude and its derivative at f1
 *            role 2  IMRPhenomPv2 frame (thetaJN, alpha0, zeta), detector projections, tiles to visit
 *            role 3  chirp mass, amplitude scale, ppE coefficients, powers of M; Pv2 precession constants
 *   phase 2  role 0  inspiral phase at f1p                 role 1  its derivative at f1p
 *            role 2  merger-ringdown phase and derivative at f2p with alpha0 = alpha1 = 0, ln(M f1p),
 *                    ln(M f2p), amplitude connection
 *            role 3  phase piece at f_ref, merger-ringdown derivative at f3 (alpha1 = 0); Pv2 Euler offsets
 *            Pv2 with shift_time: the ten knots of the t_c spline (merger-ringdown phase, alpha0 = alpha1 = 0)
 *   phase 3  role 0  the C1 phase connection (beta0, beta1, alpha0, alpha1), phase and time alignment
 *
 * The merger-ringdown phase is linear in alpha0 and alpha1: its values at f2p, at f_ref and at the spline
 * knots are evaluated before the connection fixes those two and completed afterwards (a change of the
 * rounding only).  Every other quantity follows the reference's own sequence of operations.
 */

/* what the four roles of one instance exchange */
struct SetupShared {
    SetupState S;
    double alpha1_fit;        /* alpha1 of the fits (the time shift uses it) */
    double v1, d1, v3, d3;    /* inspiral amplitude / derivative at f1, merger-ringdown at f3 */
    double alpha0, phi_aligned; /* Pv2 frame */
    double p_ins, d_ins;      /* inspiral phase and derivative at f1p */
    double q_mr, dq_mr;       /* merger-ringdown phase and derivative at f2p, alpha0 = alpha1 = 0 */
    double l1, l2;            /* ln(M f1p), ln(M f2p) */
    double ref;               /* at f_ref: inspiral phase, merger-ringdown phase (alpha0 = alpha1 = 0) or ln(M f_ref) */
    double dq_mr3;            /* merger-ringdown derivative at f3, alpha1 = 0 */
    double tq[10];            /* Pv2 t_c spline knots: merger-ringdown phase, alpha0 = alpha1 = 0 */
};

/* masses, spins and modification flags (util.cpp:996-1028); every role computes these */
__device__ __forceinline__ void gwh_setup_base(const gwh_params &p, int ppe, SetupState &S)
{
    S.m1 = p.mass1 * GWH_MSOL_SEC;
    S.m2 = p.mass2 * GWH_MSOL_SEC;
    S.chi1 = p.spin1[2];
    S.chi2 = p.spin2[2];
    S.M = S.m1 + S.m2;
    S.eta = (S.m1 * S.m2) / (S.M * S.M);
    S.chi_s = 0.5 * (S.chi1 + S.chi2);
    S.chi_a = 0.5 * (S.chi1 - S.chi2);
    S.delta = sqrt(1. - 4. * S.eta);
    S.chi_eff = (S.m1 * S.chi1 + S.m2 * S.chi2) / S.M;
    S.chi_pn = S.chi_eff - (38. * S.eta / 113.) * (2. * S.chi_s);
    /* ppE terms (waveform_generator.cpp:1296-1300 copies them for the ppE methods only) */
    S.ppe_n = 0;
    if (ppe == 1 || ppe == 2) {
        const int nm = p.ppe_theory ? gwh_theory_nterms(p.ppe_theory, p.ppe_nmod) : p.ppe_nmod;
        S.ppe_n = nm < 0 ? 0 : (nm > GWH_PPE_MAX ? GWH_PPE_MAX : nm);
    }
    S.ppe_imr = (ppe == 2);
    S.gimr = (ppe == 3);
    for (int i = 0; i < 4; i++) S.gimr_neg[i] = S.gimr ? p.gimr_dphi[i] : 0.0;
}

/* Pv2 in-plane spin parameter (IMRPhenomP.cpp:1085-1110): the given chi_p, or the one of the in-plane spins */
__device__ __forceinline__ double gwh_setup_chip(const gwh_params &p, const SetupState &S)
{
    if ((p.chip + 1) > 1e-10) return p.chip;
    const double s1x = p.spin1[0], s1y = p.spin1[1], s2x = p.spin2[0], s2y = p.spin2[1];
    const double m1_2 = S.m1 * S.m1, m2_2 = S.m2 * S.m2;
    double S1p = m1_2 * sqrt(s1y * s1y + s1x * s1x), S2p = m2_2 * sqrt(s2y * s2y + s2x * s2x);
    double A1 = 2 + (3 * S.m2) / (2 * S.m1), A2 = 2 + (3 * S.m1) / (2 * S.m2);
    double num = (A2 * S2p > A1 * S1p) ? A2 * S2p : A1 * S1p;
    double den = (S.m2 > S.m1) ? A2 * m2_2 : A1 * m1_2;
    return num / den;
}

/* role 3, phase 1: the calibration fits of the phase and of the intermediate amplitude (IMRPhenomD.cpp:852-870)
 * with the gIMR deltas (gIMRPhenomD::assign_lambda_param, gIMRPhenomD.cpp:56-87).  Returns alpha1 of the fits. */
__device__ __forceinline__ double gwh_setup_phase_fits(const gwh_params &p, const double *fits, SetupState &S)
{
    const double xfit = S.chi_pn - 1;
    S.v2 = gwh_fit(fits, 3, S.eta, xfit);
    S.sigma[0] = 0;
    for (int i = 0; i < 4; i++) S.sigma[i + 1] = gwh_fit(fits, 7 + i, S.eta, xfit);
    S.beta[0] = 0;
    for (int i = 0; i < 3; i++) S.beta[i + 1] = gwh_fit(fits, 11 + i, S.eta, xfit);
    S.alpha[0] = 0;
    for (int i = 0; i < 5; i++) S.alpha[i + 1] = gwh_fit(fits, 14 + i, S.eta, xfit);
    const double alpha1_fit = S.alpha[1];
    if (S.gimr) {
        for (int i = 2; i < 5; i++) S.sigma[i] *= (1. + p.gimr_dsigma[i - 2]);
        for (int i = 2; i < 4; i++) S.beta[i] *= (1. + p.gimr_dbeta[i - 2]);
        for (int i = 2; i < 6; i++) S.alpha[i] *= (1. + p.gimr_dalpha[i - 2]);
    }
    return alpha1_fit;
}

/* role 0, phase 1: the merger-ringdown amplitude fits gamma1..3, remnant spin and the ringdown / damping
 * frequencies (IMRPhenomD.cpp:1091-1179; IMRPhenomP.cpp:1232-1254 for the precessing remnant), transition
 * frequencies (IMRPhenomD.cpp:1312) */
__device__ __forceinline__ void gwh_setup_ringdown(const gwh_params &p, bool pv2, double chip, const double *fits,
                                                   const QnmTable &qnm, SetupState &S)
{
    const double xfit = S.chi_pn - 1;
    for (int i = 0; i < 3; i++) S.g[i] = gwh_fit(fits, 4 + i, S.eta, xfit);
    double af;
    if (pv2) {
        double Mt = S.m1 + S.m2;
        double et = S.m1 * S.m2 / (Mt * Mt);
        double qf, afp;
        if (S.m1 >= S.m2) {
            qf = S.m1 / Mt;
            afp = gwh_final_spin0815(et, S.chi1, S.chi2);
        } else {
            qf = S.m2 / Mt;
            afp = gwh_final_spin0815(et, S.chi2, S.chi1);
        }
        double Sperp = chip * qf * qf;
        af = sqrt(1.0 / (afp * afp)) * afp * sqrt(Sperp * Sperp + afp * afp);
    } else {
        af = gwh_final_spin0815(S.eta, S.chi1, S.chi2);
    }
    if (p.dep_postmerger) {
        /* IMRPhenomD.cpp:1112-1179 */
        double eta = S.eta, eta2 = eta * eta, eta3 = eta2 * eta, eta4 = eta3 * eta;
        double Ss = (S.chi1 * S.m1 * S.m1 + S.chi2 * S.m2 * S.m2) / (S.M * S.M);
        double Sred = Ss / (1. - 2. * eta);
        double Erns = 0.0559745 * eta + 0.580951 * eta2 - 0.960673 * eta3 + 3.35241 * eta4;
        double Er = Erns * (1. + Sred * (-0.00303023 - 2.00661 * eta + 7.70506 * eta2)) /
                    (1 + Sred * (-0.67144 - 1.47569 * eta + 7.30468 * eta2));
        double MWRD = 1.5251 - 1.1568 * pow(1 - af, 0.1292);
        double MWd = MWRD / (2. * (0.700 + 1.4187 * pow(1. - af, -.4990)));
        S.fRD = (1. / GWH_2PI) * MWRD / (S.M * (1. - Er));
        S.fdamp = (1. / GWH_2PI) * MWd / (S.M * (1. - Er));
    } else {
        double Er = gwh_erad0815(S.eta, S.chi1, S.chi2);
        double fr, fd;
        gwh_qnm_eval(qnm, af, fr, fd);
        S.fRD = fr / (1.0 - Er) / S.M;
        S.fdamp = fd / (1.0 - Er) / S.M;
    }
    S.f1p = 0.018 / S.M;
    S.f2p = S.fRD / 2.;
    S.f1 = 0.014 / S.M;
    {
        double g2 = S.g[1], g3 = S.g[2];
        double v = (g2 > 1) ? S.fRD + (S.fdamp * (-1.) * g3) / g2 : S.fRD + S.fdamp * g3 * (sqrt(1 - g2 * g2) - 1) / g2;
        S.f3 = fabs(v);
    }
}

/* role 1, phase 1: PN coefficients (IMRPhenomD.cpp:955-1032) with the gIMR deltas
 * (gIMRPhenomD::assign_static_pn_phase_coeff, gIMRPhenomD.cpp:89-143) */
__device__ __forceinline__ void gwh_setup_pn(const gwh_params &p, SetupState &S)
{
    double dm = S.delta, eta = S.eta, ca = S.chi_a, cs = S.chi_s;
    double eta2 = eta * eta, eta3 = eta2 * eta, cs2 = cs * cs, ca2 = ca * ca, pi2 = GWH_PI * GWH_PI;
    S.pa[0] = 1.;
    S.pa[1] = 0.;
    S.pa[2] = (-323. / 224 + 451 * eta / 168);
    S.pa[3] = (27. * dm * ca / 8 + (27. / 8 - 11. * eta / 6) * cs);
    S.pa[4] = (-27312085. / 8128512 - 1975055 * eta / 338688 + 105271 * eta2 / 24192 + (-81. / 32 + 8 * eta) * ca2 -
               81 * dm * ca * cs / 16 + (-81. / 32 + 17 * eta / 8) * cs2);
    S.pa[5] = (-85 * GWH_PI / 64 + 85 * GWH_PI * eta / 16 + dm * (285197. / 16128 - 1579 * eta / 4032) * ca +
               (285197. / 16128 - 15317 * eta / 672 - 2227 * eta2 / 1008) * cs);
    S.pa[6] = (-177520268561. / 8583708672 + (545384828789. / 5007163392 - 205 * pi2 / 48) * eta -
               3248849057 * eta2 / 178827264 + 34473079 * eta3 / 6386688 +
               (1614569. / 64512 - 1873643. * eta / 16128 + 2167 * eta2 / 42) * ca2 +
               (31 * GWH_PI / 12 - 7 * GWH_PI * eta / 3) * cs +
               (1614569. / 64512 - 61391 * eta / 1344 + 57451 * eta2 / 4032) * cs2 +
               dm * ca * (31 * GWH_PI / 12 + (1614569. / 32256 - 165961 * eta / 2688) * cs));
    for (int i = 0; i < 12; i++) S.pp[i] = 0;
    S.pp[0] = 1.;
    S.pp[2] = 3715. / 756 + 55. * eta / 9;
    S.pp[3] = -16. * GWH_PI + 113. * dm * ca / 3 + (113. / 3 - 76. * eta / 3) * cs;
    S.pp[4] = 15293365. / 508032 + 27145. * eta / 504 + 3085. * eta2 / 72 + (-405. / 8 + 200. * eta) * ca2 -
              (405. / 4) * dm * ca * cs + (-405. / 8 + 5. * eta / 2) * cs2;
    S.pp[7] = 77096675. * GWH_PI / 254016 + 378515. * GWH_PI * eta / 1512 - 74045. * GWH_PI * eta2 / 756 +
              dm * (-25150083775. / 3048192 + 26804935. * eta / 6048 - 1985. * eta2 / 48) * ca +
              (-25150083775. / 3048192 + 10566655595. * eta / 762048 - 1042165. * eta2 / 3024 + 5345. * eta3 / 36) * cs;
    S.pp[8] = gwh_pn_c8_gr(eta, dm, ca, cs);
    S.pp[9] = S.pp[8];
    S.pp[10] = 11583231236531. / 4694215680 - 6848. * GWH_GAMMA_E / 21 - 640. * pi2 / 3 +
               (-15737765635. / 3048192 + 2255. * pi2 / 12) * eta + 76055. * eta2 / 1728 - 127825. * eta3 / 1296 +
               2270. * dm * ca * GWH_PI / 3 + (2270. * GWH_PI / 3 - 520. * GWH_PI * eta) * cs -
               6848. * (GWH_LOG_64) / 63;
    S.pp[11] = 6848. / 63.;
    if (S.gimr) {
        const int direct[5] = {0, 2, 3, 4, 7};
        for (int j = 0; j < 5; j++) S.pp[direct[j]] *= (1. + p.gimr_dphi[direct[j] + 4]);
        S.pp[1] = p.gimr_dphi[1 + 4];
        S.pp[8] *= (1. + p.gimr_dphi[5 + 4]);
        S.pp[10] *= (1. + p.gimr_dphi[6 + 4]);
        S.pp[9] *= (1. + p.gimr_dphi[8 + 4]);
        S.pp[11] *= (1. + p.gimr_dphi[9 + 4]);
    }
}

/* role 3, phase 1: chirp mass and amplitude scale (util.cpp:996-1028), ppE coefficients beta (pi Mc)^(b/3), the
 * terms mapped from a theory's parameters first where ppe_theory is set.  More than GWH_PPE_MAX ppE terms, and a
 * theory that cannot be mapped, give a NaN amplitude (and so a NaN likelihood). */
__device__ __forceinline__ void gwh_setup_mass(const gwh_params &p, int ppe, SetupState &S)
{
    S.Mc = pow(S.m1 * S.m2, 0.6) / pow(S.M, 0.2);
    double DL = p.DL * (GWH_MPC_SI / GWH_C_SI);
    double kA = p.sky_average ? sqrt(GWH_PI / 30.) : sqrt(GWH_PI * 40. / 192.);
    S.A0 = kA * S.Mc * S.Mc / DL * pow(GWH_PI * S.Mc, -7. / 6.);
    for (int i = 0; i < GWH_PPE_MAX; i++) S.ppe_e[i] = S.ppe_c[i] = 0.0;
    if (ppe == 1 || ppe == 2) {
        double b[GWH_PPE_MAX], beta[GWH_PPE_MAX];
        bool bad = p.ppe_nmod > GWH_PPE_MAX;
        for (int i = 0; i < GWH_PPE_MAX; i++) {
            b[i] = p.ppe_b[i];
            beta[i] = p.ppe_beta[i];
        }
        if (p.ppe_theory) bad |= gwh_theory_terms(p, S.M, S.Mc, S.eta, b, beta) < 0;
        if (bad) S.A0 = __longlong_as_double(0x7ff8000000000000LL); /* NaN amplitude */
        const double lpm = log(GWH_PI * S.Mc);
        for (int i = 0; i < GWH_PPE_MAX; i++) {
            S.ppe_e[i] = i < S.ppe_n ? b[i] / 3.0 : 0.0;
            S.ppe_c[i] = i < S.ppe_n ? beta[i] * exp(S.ppe_e[i] * lpm) : 0.0;
        }
    }
}

/* role 2, phase 1: the Pv2 frame (IMRPhenomP.cpp:821-938, :1081-1211): thetaJN, alpha0, zeta and the
 * aligned-spin phase offset -phiJ */
struct Pv2Frame {
    double thetaJN, alpha0, zeta, phi_aligned;
};
/* rotation by an angle given as its cosine and sine (the frame applies each of its three rotations several
 * times: one sincos per angle) */
__device__ __forceinline__ void gwh_rotz_cs(double c, double s, double &x, double &y)
{
    double t1 = x * c - y * s, t2 = x * s + y * c;
    x = t1;
    y = t2;
}
__device__ __forceinline__ void gwh_roty_cs(double c, double s, double &x, double &z)
{
    double t1 = x * c + z * s, t2 = -x * s + z * c;
    x = t1;
    z = t2;
}

__device__ __forceinline__ Pv2Frame gwh_setup_frame(const gwh_params &p, const SetupState &S, double chip)
{
    const double s1x = p.spin1[0], s1y = p.spin1[1], s2x = p.spin2[0], s2y = p.spin2[1];
    const bool reduced = (p.chip + 1) > 1e-10;
    const double m1_2 = S.m1 * S.m1, m2_2 = S.m2 * S.m2;
    double xr = cbrt(GWH_PI * S.M * p.f_ref);
    double L0 = S.M * S.M * gwh_l2pn(S.eta, xr * xr);
    double J0x, J0y;
    if (reduced) {
        double sp, cp;
        sincos(p.phip, &sp, &cp);
        J0x = m1_2 * chip * cp;
        J0y = m1_2 * chip * sp;
    } else {
        J0x = m1_2 * s1x + m2_2 * s2x;
        J0y = m1_2 * s1y + m2_2 * s2y;
    }
    double J0z = L0 + m1_2 * S.chi1 + m2_2 * S.chi2;
    double J0 = sqrt(J0x * J0x + J0y * J0y + J0z * J0z);
    double thetaJ = acos(J0z / J0), phiJ = atan2(J0y, J0x);
    double sJ, cJ, sP, cP; /* rotations by -thetaJ and -phiJ */
    sincos(thetaJ, &sJ, &cJ);
    sincos(phiJ, &sP, &cP);
    sJ = -sJ;
    sP = -sP;
    Pv2Frame F;
    F.phi_aligned = -phiJ;
    double si, ci, sr, cr;
    sincos(p.incl, &si, &ci);
    sincos(GWH_PI / 2. - p.phiRef, &sr, &cr);
    double Nx = si * cr, Ny = si * sr, Nz = ci;
    double tx = Nx, ty = Ny, tz = Nz;
    gwh_rotz_cs(cP, sP, tx, ty);
    gwh_roty_cs(cJ, sJ, tx, tz);
    double kappa = -atan2(ty, tx);
    double sK, cK;
    sincos(kappa, &sK, &cK);
    tx = 0.;
    ty = 0.;
    tz = 1.;
    gwh_rotz_cs(cP, sP, tx, ty);
    gwh_roty_cs(cJ, sJ, tx, tz);
    gwh_rotz_cs(cK, sK, tx, ty);
    F.alpha0 = atan2(ty, tx);
    tx = Nx;
    ty = Ny;
    tz = Nz;
    gwh_rotz_cs(cP, sP, tx, ty);
    gwh_roty_cs(cJ, sJ, tx, tz);
    gwh_rotz_cs(cK, sK, tx, ty);
    double NxJ = tx, NzJ = tz;
    F.thetaJN = acos(NzJ);
    double sph, cph;
    sincos(p.phiRef, &sph, &cph);
    tx = -ci * sph;
    ty = -ci * cph;
    tz = si;
    gwh_rotz_cs(cP, sP, tx, ty);
    gwh_roty_cs(cJ, sJ, tx, tz);
    gwh_rotz_cs(cK, sK, tx, ty);
    F.zeta = atan2(tx * NzJ - tz * NxJ, -ty);
    return F;
}

/* role 2, phase 1: polarisation constants and, per detector, the projection constants and the time-of-flight
 * phase slope (waveform_util.cpp:152-184, detector_util.cpp:768-790 and :900-1016) */
__device__ __forceinline__ void gwh_setup_detectors(const gwh_params &p, bool pv2, double zeta, const DetGeom *geo,
                                                    const int *dets, int ndet, InstCoef &C)
{
    /* the detectors' geometry is loaded before the trigonometry, which then covers the load latency */
    DetGeom gd[GWH_MAXDET];
#pragma unroll
    for (int d = 0; d < GWH_MAXDET; d++)
        if (d < ndet) gd[d] = geo[dets[d]];
    double ci = cos(p.incl);
    double kplus = .5 * (1. + ci * ci);
    double c2z = cos(2. * zeta), s2z = sin(2. * zeta);
    const DetGeom &g0 = gd[0];
    C.wf_a = pv2 ? c2z : kplus;
    C.wf_b = pv2 ? s2z : ci;
    /* polarisation basis (shared by every detector) */
    const double gha = p.gmst - p.RA;
    double cg, sg, cd, sd, cp, sp;
    sincos(gha, &sg, &cg);
    sincos(p.DEC, &sd, &cd);
    sincos(p.psi, &sp, &cp);
    const double X[3] = {-cp * sg - sp * cg * sd, -cp * cg + sp * sg * sd, sp * cd};
    const double Y[3] = {sp * sg - cp * cg * sd, sp * cg + cp * sg * sd, cp * cd};
    /* propagation direction for dt = (r0 - rd).n / c (hour angle gmst - RA) */
    const double e0 = cd * cg, e1 = cd * -sg, e2 = sd;
#pragma unroll
    for (int d = 0; d < GWH_MAXDET; d++) {
        double k0 = 0, k1 = 0, k2 = 0, k3 = 0;
        if (d < ndet) {
            const DetGeom &g = gd[d];
            double fp = 0, fx = 0;
            for (int i = 0; i < 3; i++) {
                double DX = g.tensor[3 * i] * X[0] + g.tensor[3 * i + 1] * X[1] + g.tensor[3 * i + 2] * X[2];
                double DY = g.tensor[3 * i] * Y[0] + g.tensor[3 * i + 1] * Y[1] + g.tensor[3 * i + 2] * Y[2];
                fp += X[i] * DX - Y[i] * DY;
                fx += X[i] * DY + Y[i] * DX;
            }
            fp *= g.gfac;
            fx *= g.gfac;
            double dt =
                ((g0.loc[0] - g.loc[0]) * e0 + (g0.loc[1] - g.loc[1]) * e1 + (g0.loc[2] - g.loc[2]) * e2) / GWH_C_SI;
            if (!pv2) {
                k0 = fp * kplus;
                k1 = -fx * ci;
                k2 = k0 * k0 + k1 * k1;
            } else {
                k0 = fp * c2z - fx * s2z;
                k1 = fp * s2z + fx * c2z;
            }
            k3 = GWH_2PI * dt;
        }
        C.det[d][0] = k0;
        C.det[d][1] = k1;
        C.det[d][2] = k2;
        C.det[d][3] = k3;
    }
}

/* role 2, phase 2: amplitude C1 connection, the Hermite quartic through (f1, v1, v1'), (f2, v2), (f3, v3, v3')
 * with f2 = (f1 + f3) / 2 (IMRPhenomD.cpp:1679-1702), as coefficients of powers of f */
__device__ __forceinline__ void gwh_setup_amp_connection(double M, double f1, double f3, double v1, double d1,
                                                         double v2, double v3, double d3, double *out)
{
    double f2 = (f1 + f3) / 2.;
    double q12 = (v2 - v1) / (f2 - f1), q23 = (v3 - v2) / (f3 - f2);
    double q012 = (q12 - d1) / (f2 - f1), q123 = (q23 - q12) / (f3 - f1), q234 = (d3 - q23) / (f3 - f2);
    double q0123 = (q123 - q012) / (f3 - f1), q1234 = (q234 - q123) / (f3 - f1);
    double q01234 = (q1234 - q0123) / (f3 - f1);
    double a[5] = {v1, d1, q012, q0123, q01234};
    double z[4] = {f1, f1, f2, f3};
    double poly[5] = {a[4], 0, 0, 0, 0};
#pragma unroll
    for (int k = 3; k >= 0; k--) {
        double nw[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4 - k; j++) {
            nw[j + 1] += poly[j];
            nw[j] -= z[k] * poly[j];
        }
        nw[0] += a[k];
#pragma unroll
        for (int j = 0; j < 5; j++) poly[j] = nw[j];
    }
    double Mk = 1.0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        out[k] = poly[k] / Mk;
        Mk *= M;
    }
}

/* phase 3: derivative at fRD of the natural cubic spline through the 10 knots (f_j, ph_j) on [0.8, 1.2] fRD
 * (IMRPhenomP.cpp:550-606), divided by 2 pi; 0 when the knots collapse */
__device__ __forceinline__ double gwh_setup_tcorr(double fRD, const double *ph)
{
    const int n = 10;
    double start = .8 * fRD, stop = 1.2 * fRD;
    double df = (stop - start) / (n - 1);
    if (!(df > 0)) return 0;
    double fs[10], c[10], dg[10], rh[10], of[10];
#pragma unroll
    for (int j = 0; j < n; j++) fs[j] = start + j * df;
    const int m = n - 2;
#pragma unroll
    for (int i = 0; i < m; i++) {
        double h0 = fs[i + 1] - fs[i], h1 = fs[i + 2] - fs[i + 1];
        dg[i] = 2.0 * (h0 + h1);
        of[i] = h1;
        rh[i] = 3.0 * ((ph[i + 2] - ph[i + 1]) / h1 - (ph[i + 1] - ph[i]) / h0);
    }
#pragma unroll
    for (int i = 1; i < m; i++) {
        double w = of[i - 1] / dg[i - 1];
        dg[i] -= w * of[i - 1];
        rh[i] -= w * rh[i - 1];
    }
    c[0] = 0;
    c[n - 1] = 0;
    c[m] = rh[m - 1] / dg[m - 1];
#pragma unroll
    for (int i = m - 2; i >= 0; i--) c[i + 1] = (rh[i] - of[i] * c[i + 2]) / dg[i];
    /* interval holding fRD (the bisection's answer, found by an unrolled scan so that every array index is a
     * compile-time constant and the arrays stay in registers) */
    double x0 = fs[0], x1 = fs[1], y0 = ph[0], y1 = ph[1], c0 = c[0], c1 = c[1];
#pragma unroll
    for (int j = 1; j < n - 1; j++)
        if (fs[j] <= fRD) {
            x0 = fs[j];
            x1 = fs[j + 1];
            y0 = ph[j];
            y1 = ph[j + 1];
            c0 = c[j];
            c1 = c[j + 1];
        }
    double h = x1 - x0;
    double b = (y1 - y0) / h - h * (c1 + 2.0 * c0) / 3.0;
    double dd = (c1 - c0) / (3.0 * h);
    double dx = fRD - x0;
    return (b + dx * (2.0 * c0 + 3.0 * dd * dx)) / GWH_2PI;
}

/* the t_c spline's knot j: 0.8 fRD + j (0.4 fRD / 9) */
__device__ __forceinline__ double gwh_setup_knot(double fRD, int j)
{
    double start = .8 * fRD, stop = 1.2 * fRD;
    return start + j * ((stop - start) / 9);
}

/* ------------------------------------------------------------------------- */
/* per-bin FP64 evaluation                                                    */
/* ------------------------------------------------------------------------- */

/* IMRPhenomD amplitude A(f) and phase phi(f) = Psi(f) - tc_total (f - f_ref) - phic
 * (IMRPhenomD.cpp:471-501); caller guarantees 0 < f <= fcut. */
/* gIMR inspiral phase terms at one bin from 1/v, v = (pi M f)^(1/3) */
__device__ __forceinline__ double gwh_gimr_bin(const InstCoef &C, double iv)
{
    const double iv2 = iv * iv, iv4 = iv2 * iv2, iv6 = iv4 * iv2;
    return C.K * (C.cp[1] * iv4 + iv6 * (C.gimr_neg[3] + iv * (C.gimr_neg[2] + iv * (C.gimr_neg[1] +
                                                                                      iv * C.gimr_neg[0]))));
}

/* ppE phase terms at one bin from its tabulated log f: sum_i beta_i (pi Mc f)^(b_i/3) */
__device__ __forceinline__ double gwh_ppe_bin(const InstCoef &C, double lf)
{
    double t = 0;
    for (int i = 0; i < C.ppe_n; i++) t += C.ppe_c[i] * exp(C.ppe_e[i] * lf);
    return t;
}

template <bool FP32_EXP = false>
__device__ __forceinline__ void gwh_phenomd_fp64(const InstCoef &C, double f, const BinPow &bp, double &A, double &phi,
                                                 double &u_out, double &lx_out, const double2 *tab = gwh_turn_table,
                                                 const double *etab = gwh_exp2_table)
{
    double Mf = C.M * f;
    double u = C.M13 * bp.f13;      /* (Mf)^{1/3} */
    double inv76 = C.Mm76 * bp.fm76; /* (Mf)^{-7/6} */
    double ar;
    if (f < C.f1) {
        double u2 = u * u, u3 = u2 * u;
        double pn = C.ca[0] + u2 * (C.ca[2] + u * (C.ca[3] + u * (C.ca[4] + u * (C.ca[5] + u * C.ca[6]))));
        double nr = u3 * u3 * u * (C.rho[0] + u * (C.rho[1] + u * C.rho[2]));
        ar = pn + nr;
    } else if (f > C.f3) {
        double dl = f - C.fRD;
        ar = C.mr_k * (FP32_EXP ? exp_fp32(-C.mr_g2 * dl) : gwh_exp_table(-C.mr_g2 * dl, etab)) * rcp_newton(dl * dl + C.mr_w2);
    } else {
        ar = C.delta[0] + Mf * (C.delta[1] + Mf * (C.delta[2] + Mf * (C.delta[3] + Mf * C.delta[4])));
    }
    A = C.amp_pref * ar * inv76;
    double psi;
    double lx = 0;
    if (f < C.f1p) {
        double v = u * 1.4645918875615231; /* pi^{1/3} */
        lx = C.lgPiM + bp.lf;             /* log(pi M f) */
        double c5 = C.c8 + lx * C.c9;
        double c6 = C.c10 - C.c11 * lx;
        double v2 = v * v;
        double poly = C.cp[0] + v2 * (C.cp[2] + v * (C.cp[3] + v * (C.cp[4] + v * (c5 + v * (c6 + v * C.cp[7])))));
        double Mf43 = Mf * u;
        psi = -GWH_PI / 4. + C.K * poly * rcp_newton(v2 * v2 * v) +
              (C.sig[0] * Mf + C.sig[1] * Mf43 + C.sig[2] * Mf43 * u + C.sig[3] * Mf * Mf);
        if (C.gimr) psi += gwh_gimr_bin(C, rcp_newton(v));
    } else if (f > C.f2p) {
        psi = C.alpha[0] + C.alpha[1] * Mf - C.alpha[2] * rcp_newton(Mf) + C.alpha[3] * (C.M34 * bp.f34) +
              C.alpha[4] * gwh_atan_seeded((f - C.alpha[5]) * C.inv_fdamp, tab);
    } else {
        psi = C.beta[0] + C.beta[1] * Mf + C.beta[2] * (C.lgM + bp.lf) - C.beta[3] * rcp_newton(Mf * Mf * Mf);
    }
    if (C.ppe_n > 0 && (C.ppe_imr || f < C.f1p)) psi += gwh_ppe_bin(C, bp.lf);
    phi = psi - (C.tc_total * (f - C.f_ref) + C.phic);
    u_out = u;
    lx_out = lx;
}

/* IMRPhenomPv2 twist-up at one bin (IMRPhenomP.cpp:290-361): returns the common amplitude
 * (already including the 1/2), the total phase and the complex h+ / hx twist-up factors. */
__device__ __forceinline__ void gwh_pv2_fp64(const InstCoef &C, double f, const BinPow &bp, double &A, double &Phi,
                                             double &hpr, double &hpi, double &hcr, double &hci,
                                             const double2 *tab = gwh_turn_table, const double *etab = gwh_exp2_table)
{
    double u, lx;
    gwh_phenomd_fp64(C, f, bp, A, Phi, u, lx, tab, etab);
    /* gwh_phenomd_fp64 already applied -tc (f - f_ref) - phic; Pv2 adds 2 eps and the t_c correction */
    double v = u * 1.4645918875615231;
    double logw = C.lgPiM + bp.lf; /* log(pi M f) */
    double x = v * v;
    double L = C.eta * (1.0 + C.L_c1 * x + C.L_c2 * x * x) * rcp_newton(v);
    double s = C.SP * rcp_newton(L + C.SL);
    double cb = rsqrt_seeded(1.0 + s * s);
    /* cos and sin of beta/2: 1 - cos b = s^2 cos^2 b / (1 + cos b), so sin(b/2) = |s| cos b / (2 cos(b/2)), free of
       the cancellation of sqrt((1 - cos b) / 2) for small s and one square root fewer */
    double cbh = sqrt((1.0 + cb) / 2.0), sbh = fabs(s) * cb * (0.5 * rcp_newton(cbh));
    double c2 = cbh * cbh, s2 = sbh * sbh, c3 = c2 * cbh, s3 = s2 * sbh;
    double d2[5] = {s3 * sbh, 2 * cbh * s3, GWH_SQRT_6 * s2 * c2, 2 * c3 * sbh, c3 * cbh};
    double alpha = gwh_euler_eval(C.a_c, v, logw) + C.alpha_add;
    double eps = gwh_euler_eval(C.e_c, v, logw) + C.eps_add;
    double sa, ca;
    gwh_sincos_tab(alpha, sa, ca, tab);
    /* e^{i m alpha}, m=-2..2 */
    double er[5], ei[5];
    er[2] = 1;
    ei[2] = 0;
    er[3] = ca;
    ei[3] = sa;
    er[1] = ca;
    ei[1] = -sa;
    er[4] = ca * ca - sa * sa;
    ei[4] = 2 * ca * sa;
    er[0] = er[4];
    ei[0] = -ei[4];
    hpr = hpi = hcr = hci = 0;
#pragma unroll
    for (int m = -2; m <= 2; m++) {
        double dm2 = (m == -2) ? d2[4] : (m == -1) ? -d2[3] : (m == 0) ? d2[2] : (m == 1) ? -d2[1] : d2[0];
        double y = C.ylm[m + 2];
        double t2r = er[-m + 2] * dm2 * y, t2i = ei[-m + 2] * dm2 * y;
        double tmr = er[m + 2] * d2[m + 2] * y, tmi = ei[m + 2] * d2[m + 2] * y;
        hpr += t2r + tmr;
        hpi += t2i + tmi;
        hcr -= t2i - tmi;
        hci += t2r - tmr;
    }
    Phi = Phi + 2. * eps + C.t_corr_2pi * f;
}

#endif
