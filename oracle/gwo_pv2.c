/*
 * IMRPhenomPv2 oracle (TEST INFRASTRUCTURE): FP64 restatement of
 * src/IMRPhenomP.cpp (twist-up of the IMRPhenomD (2,2) model with
 * single-spin precession Euler angles).
 */
#include <stdlib.h>
#include <string.h>
#include "gwo_internal.h"

/* IMRPhenomP.h:221-231 rotations about z and y */
static void rotz(double a, double *x, double *y)
{
    double t1 = *x * cos(a) - *y * sin(a);
    double t2 = *x * sin(a) + *y * cos(a);
    *x = t1;
    *y = t2;
}
static void roty(double a, double *x, double *z)
{
    double t1 = *x * cos(a) + *z * sin(a);
    double t2 = -*x * sin(a) + *z * cos(a);
    *x = t1;
    *z = t2;
}

/* IMRPhenomP.cpp:1213-1220: 2PN orbital angular momentum / M^2 */
static double L2PN(double eta, const gwo_pows *pw)
{
    double x = pw->MF2third * pw->PI2third;
    double x2 = x * x, eta2 = eta * eta;
    return (eta * (1.0 + (1.5 + eta / 6.0) * x + (3.375 - (19.0 * eta) / 8. - eta2 / 24.0) * x2)) / sqrt(x);
}

/* IMRPhenomP.cpp:821-938 (reduced=1, from chip/phip) and :1081-1211 (reduced=0, from spin vectors) */
void gwo_pv2_param_transform(gwo_src *s, int reduced)
{
    double chi1_l = s->spin1z, chi2_l = s->spin2z;
    double q = s->mass1 / s->mass2;
    double chi_eff = (s->mass1 * chi1_l + s->mass2 * chi2_l) / s->M;
    s->chil = (1.0 + q) / q * chi_eff;
    double m1_2 = s->mass1 * s->mass1, m2_2 = s->mass2 * s->mass2;
    if (!reduced) {
        double S1_perp = m1_2 * sqrt(s->spin1y * s->spin1y + s->spin1x * s->spin1x);
        double S2_perp = m2_2 * sqrt(s->spin2y * s->spin2y + s->spin2x * s->spin2x);
        double A1 = 2 + (3 * s->mass2) / (2 * s->mass1);
        double A2 = 2 + (3 * s->mass1) / (2 * s->mass2);
        double ASp1 = A1 * S1_perp, ASp2 = A2 * S2_perp;
        double num = (ASp2 > ASp1) ? ASp2 : ASp1;
        double denom = (s->mass2 > s->mass1) ? A2 * m2_2 : A1 * m1_2;
        s->chip = num / denom;
    }
    double m1 = q / (1 + q), m2 = 1. / (1 + q);
    s->SP = s->chip * m1 * m1;
    s->SL = chi1_l * m1 * m1 + chi2_l * m2 * m2;

    gwo_pows pw;
    gwo_precalc_powers_pi(&pw);
    gwo_precalc_powers_ins(s->f_ref, s->M, &pw);
    double L0 = s->M * s->M * L2PN(s->eta, &pw);
    double J0x, J0y;
    if (reduced) {
        J0x = m1_2 * s->chip * cos(s->phip);
        J0y = m1_2 * s->chip * sin(s->phip);
    } else {
        J0x = m1_2 * s->spin1x + m2_2 * s->spin2x;
        J0y = m1_2 * s->spin1y + m2_2 * s->spin2y;
    }
    double J0z = L0 + m1_2 * s->spin1z + m2_2 * s->spin2z;
    double J0 = sqrt(J0x * J0x + J0y * J0y + J0z * J0z);
    double thetaJ = acos(J0z / J0);
    double phiJ = atan2(J0y, J0x);
    s->phi_aligned = -phiJ;

    double incl = s->incl_angle, phiRef = s->phiRef;
    double Nx = sin(incl) * cos(M_PI / 2. - phiRef);
    double Ny = sin(incl) * sin(M_PI / 2. - phiRef);
    double Nz = cos(incl);
    double tx = Nx, ty = Ny, tz = Nz;
    rotz(-phiJ, &tx, &ty);
    roty(-thetaJ, &tx, &tz);
    double kappa = -atan2(ty, tx);

    tx = 0.;
    ty = 0.;
    tz = 1.;
    rotz(-phiJ, &tx, &ty);
    roty(-thetaJ, &tx, &tz);
    rotz(kappa, &tx, &ty);
    s->alpha0 = atan2(ty, tx);

    tx = Nx;
    ty = Ny;
    tz = Nz;
    rotz(-phiJ, &tx, &ty);
    roty(-thetaJ, &tx, &tz);
    rotz(kappa, &tx, &ty);
    double Nx_Jf = tx, Nz_Jf = tz;
    s->thetaJN = acos(Nz_Jf);

    double Xx = -cos(incl) * sin(phiRef), Xy = -cos(incl) * cos(phiRef), Xz = sin(incl);
    tx = Xx;
    ty = Xy;
    tz = Xz;
    rotz(-phiJ, &tx, &ty);
    roty(-thetaJ, &tx, &tz);
    rotz(kappa, &tx, &ty);
    /* P = (0,-1,0), Q = (Nz, 0, -Nx) in the J frame (Arun et al. triad) */
    double XdotP = -ty;
    double XdotQ = tx * Nz_Jf - tz * Nx_Jf;
    s->zeta_polariz = atan2(XdotQ, XdotP);
}

typedef struct { double c1, c2, c3, c4, c5; } euler_coeffs;

/* IMRPhenomP.cpp:64-144: NNLO single-spin Euler-angle coefficients (alpha and epsilon) */
static void euler_coeffs_calc(euler_coeffs *a, euler_coeffs *e, const gwo_src *s)
{
    double m2 = s->q / (1. + s->q), m1 = 1. / (1. + s->q);
    double dm = m1 - m2, mtot = 1.;
    double eta = m1 * m2, eta2 = eta * eta, eta3 = eta2 * eta, eta4 = eta3 * eta;
    double mtot2 = mtot * mtot, mtot4 = mtot2 * mtot2, mtot6 = mtot4 * mtot2, mtot8 = mtot6 * mtot2;
    double chil = s->chil, chil2 = chil * chil;
    double chip2 = s->chip * s->chip, chip4 = chip2 * chip2;
    double dm2 = dm * dm, dm3 = dm2 * dm;
    double m2_2 = m2 * m2, m2_3 = m2_2 * m2, m2_4 = m2_3 * m2, m2_5 = m2_4 * m2, m2_6 = m2_5 * m2,
           m2_7 = m2_6 * m2, m2_8 = m2_7 * m2;

    a->c1 = (-0.18229166666666666 - (5 * dm) / (64. * m2));
    a->c2 = ((-15 * dm * m2 * chil) / (128. * mtot2 * eta) - (35 * m2_2 * chil) / (128. * mtot2 * eta));
    a->c3 = (-1.7952473958333333 - (4555 * dm) / (7168. * m2) - (15 * chip2 * dm * m2_3) / (128. * mtot4 * eta2) -
             (35 * chip2 * m2_4) / (128. * mtot4 * eta2) - (515 * eta) / 384. - (15 * dm2 * eta) / (256. * m2_2) -
             (175 * dm * eta) / (256. * m2));
    a->c4 = -(35 * M_PI) / 48. - (5 * dm * M_PI) / (16. * m2) + (5 * dm2 * chil) / (16. * mtot2) +
            (5 * dm * m2 * chil) / (3. * mtot2) + (2545 * m2_2 * chil) / (1152. * mtot2) -
            (5 * chip2 * dm * m2_5 * chil) / (128. * mtot6 * eta3) - (35 * chip2 * m2_6 * chil) / (384. * mtot6 * eta3) +
            (2035 * dm * m2 * chil) / (21504. * mtot2 * eta) + (2995 * m2_2 * chil) / (9216. * mtot2 * eta);
    a->c5 = (4.318908476114694 + (27895885 * dm) / (2.1676032e7 * m2) - (15 * chip4 * dm * m2_7) / (512. * mtot8 * eta4) -
             (35 * chip4 * m2_8) / (512. * mtot8 * eta4) - (485 * chip2 * dm * m2_3) / (14336. * mtot4 * eta2) +
             (475 * chip2 * m2_4) / (6144. * mtot4 * eta2) + (15 * chip2 * dm2 * m2_2) / (256. * mtot4 * eta) +
             (145 * chip2 * dm * m2_3) / (512. * mtot4 * eta) + (575 * chip2 * m2_4) / (1536. * mtot4 * eta) +
             (39695 * eta) / 86016. + (1615 * dm2 * eta) / (28672. * m2_2) - (265 * dm * eta) / (14336. * m2) +
             (955 * eta2) / 576. + (15 * dm3 * eta2) / (1024. * m2_3) + (35 * dm2 * eta2) / (256. * m2_2) +
             (2725 * dm * eta2) / (3072. * m2) - (15 * dm * m2 * M_PI * chil) / (16. * mtot2 * eta) -
             (35 * m2_2 * M_PI * chil) / (16. * mtot2 * eta) + (15 * chip2 * dm * m2_7 * chil2) / (128. * mtot8 * eta4) +
             (35 * chip2 * m2_8 * chil2) / (128. * mtot8 * eta4) + (375 * dm2 * m2_2 * chil2) / (256. * mtot4 * eta) +
             (1815 * dm * m2_3 * chil2) / (256. * mtot4 * eta) + (1645 * m2_4 * chil2) / (192. * mtot4 * eta));

    e->c1 = (-0.18229166666666666 - (5 * dm) / (64. * m2));
    e->c2 = ((-15 * dm * m2 * chil) / (128. * mtot2 * eta) - (35 * m2_2 * chil) / (128. * mtot2 * eta));
    e->c3 = (-1.7952473958333333 - (4555 * dm) / (7168. * m2) - (515 * eta) / 384. - (15 * dm2 * eta) / (256. * m2_2) -
             (175 * dm * eta) / (256. * m2));
    e->c4 = -(35 * M_PI) / 48. - (5 * dm * M_PI) / (16. * m2) + (5 * dm2 * chil) / (16. * mtot2) +
            (5 * dm * m2 * chil) / (3. * mtot2) + (2545 * m2_2 * chil) / (1152. * mtot2) +
            (2035 * dm * m2 * chil) / (21504. * mtot2 * eta) + (2995 * m2_2 * chil) / (9216. * mtot2 * eta);
    e->c5 = (4.318908476114694 + (27895885 * dm) / (2.1676032e7 * m2) + (39695 * eta) / 86016. +
             (1615 * dm2 * eta) / (28672. * m2_2) - (265 * dm * eta) / (14336. * m2) + (955 * eta2) / 576. +
             (15 * dm3 * eta2) / (1024. * m2_3) + (35 * dm2 * eta2) / (256. * m2_2) + (2725 * dm * eta2) / (3072. * m2) -
             (15 * dm * m2 * M_PI * chil) / (16. * mtot2 * eta) - (35 * m2_2 * M_PI * chil) / (16. * mtot2 * eta) +
             (375 * dm2 * m2_2 * chil2) / (256. * mtot4 * eta) + (1815 * dm * m2_3 * chil2) / (256. * mtot4 * eta) +
             (1645 * m2_4 * chil2) / (192. * mtot4 * eta));
}

/* IMRPhenomP.cpp:716-735 */
static void euler_angles(double *alpha, double *epsilon, const gwo_pows *pw, const euler_coeffs *a, const euler_coeffs *e)
{
    double oc = pw->MFthird * pw->PIthird;
    double oc2 = oc * oc;
    double logomega = log(oc2 * oc);
    double omega = oc2 * oc;
    *alpha = (a->c1 / omega + a->c2 / oc2 + a->c3 / oc + a->c4 * logomega + a->c5 * oc);
    *epsilon = (e->c1 / omega + e->c2 / oc2 + e->c3 / oc + e->c4 * logomega + e->c5 * oc);
}

/* IMRPhenomP.cpp:664-693: Wigner-d^2_{m,2}(beta) (d2) and d^2_{m,-2} (dm2) */
static void wigner_d(double d2[5], double dm2[5], const gwo_pows *pw, const gwo_src *s)
{
    double L = L2PN(s->eta, pw);
    double sv = s->SP / (L + s->SL);
    double s_2 = sv * sv;
    double cos_beta = 1. / sqrt(1.0 + s_2);
    double cbh = sqrt((1.0 + cos_beta) / 2.0);
    double sbh = sqrt((1.0 - cos_beta) / 2.0);
    double c2 = cbh * cbh, s2 = sbh * sbh;
    double c3 = c2 * cbh, s3 = s2 * sbh;
    double c4 = c3 * cbh, s4 = s3 * sbh;
    d2[0] = s4;
    d2[1] = 2 * cbh * s3;
    d2[2] = GWO_SQRT_6 * s2 * c2;
    d2[3] = 2 * c3 * sbh;
    d2[4] = c4;
    dm2[0] = d2[4];
    dm2[1] = -d2[3];
    dm2[2] = d2[2];
    dm2[3] = -d2[1];
    dm2[4] = d2[0];
}

/* util.cpp:1953-2000 (s=-2, l=2, phi=0): real spin-weighted harmonics -2Y_2m(theta) */
static void ylm2(double theta, double y[5])
{
    y[0] = sqrt(5.0 / (64.0 * M_PI)) * (1.0 - cos(theta)) * (1.0 - cos(theta));
    y[1] = sqrt(5.0 / (16.0 * M_PI)) * sin(theta) * (1.0 - cos(theta));
    y[2] = sqrt(15.0 / (32.0 * M_PI)) * sin(theta) * sin(theta);
    y[3] = sqrt(5.0 / (16.0 * M_PI)) * sin(theta) * (1.0 + cos(theta));
    y[4] = sqrt(5.0 / (64.0 * M_PI)) * (1.0 + cos(theta)) * (1.0 + cos(theta));
}

/* IMRPhenomP.cpp:695-715: h+ and hx factors of the twisted-up (2,+-2) modes */
static void twistup(double alpha, const double d2[5], const double dm2[5], const double y[5], double *hp, double *hc)
{
    double ca = cos(alpha), sa = sin(alpha);
    /* e^{i m alpha}, m = -2..2 */
    double er[5], ei[5];
    double den = ca * ca + sa * sa;
    double mr = ca / den, mi = -sa / den; /* 1/e^{i alpha} */
    er[3] = ca;
    ei[3] = sa;
    er[1] = mr;
    ei[1] = mi;
    er[2] = 1.0;
    ei[2] = 0.0;
    gwo_cmul(ca, sa, ca, sa, &er[4], &ei[4]);
    gwo_cmul(mr, mi, mr, mi, &er[0], &ei[0]);
    hp[0] = hp[1] = hc[0] = hc[1] = 0.0;
    for (int m = -2; m <= 2; m++) {
        double t2r = er[-m + 2] * dm2[m + 2] * y[m + 2], t2i = ei[-m + 2] * dm2[m + 2] * y[m + 2];
        double tmr = er[m + 2] * d2[m + 2] * y[m + 2], tmi = ei[m + 2] * d2[m + 2] * y[m + 2];
        hp[0] += t2r + tmr;
        hp[1] += t2i + tmi;
        /* i*(T2m - Tm2m) */
        hc[0] += -(t2i - tmi);
        hc[1] += (t2r - tmr);
    }
}

/* IMRPhenomP.cpp:550-606: t_c correction from the derivative of a 10-knot natural
 * cubic spline of -Psi_MR on [0.8, 1.2] f_RD, evaluated at f_RD */
static double time_shift(gwo_src *s, gwo_phenomd_state *st, gwo_pows *pw)
{
    double M = s->M;
    const int n = 10;
    double f_final = s->fRD;
    double fs[10], ph[10];
    double start = .8 * (f_final), stop = 1.2 * (f_final);
    double df = (stop - start) / (n - 1);
    if (isnan(df) || df <= 0) return 0;
    for (int j = 0; j < n; j++) fs[j] = start + j * df;
    for (int j = 0; j < n; j++) {
        double f = fs[j];
        if (f < s->f1_phase) {
            gwo_precalc_powers_ins(f, M, pw);
        } else {
            pw->MFsixth = pow(M * f, 1. / 6);
        }
        ph[j] = -gwo_build_phase(f, s, st, pw);
    }
    return gwo_spline_eval(fs, ph, n, f_final, 1) / (2. * M_PI);
}

/* IMRPhenomP.cpp:194-369 construct_waveform */
int gwo_pv2_waveform(const double *f, int n, double *hp, double *hc, gwo_src *s)
{
    double y[5];
    ylm2(s->thetaJN, y);
    gwo_phenomd_state st;
    double M = s->M;
    int save_shift_time = s->shift_time, save_shift_phase = s->shift_phase;
    /* the PhenomD phase/time shifts are not used by the Pv2 twist-up */
    s->shift_time = 0;
    s->shift_phase = 0;
    gwo_phenomd_prepare(s, &st);
    s->shift_time = save_shift_time;
    s->shift_phase = save_shift_phase;

    double f_ref = s->f_ref;
    double phic = 2 * s->phi_aligned;
    double tc = 2 * M_PI * s->tc;
    double A0 = s->A0 * pow(M, 7. / 6.) / (2. * sqrt(5. / (64. * M_PI)));
    s->q = s->mass1 / s->mass2;
    euler_coeffs ac, ec;
    euler_coeffs_calc(&ac, &ec, s);

    gwo_pows pw;
    gwo_precalc_powers_pi(&pw);
    double *amp_vec = (double *)malloc(sizeof(double) * n);
    double *phase_vec = (double *)malloc(sizeof(double) * n);
    double *hpf = (double *)malloc(sizeof(double) * 2 * n);
    double *hcf = (double *)malloc(sizeof(double) * 2 * n);

    pw.MFthird = pow(s->M * s->f_ref, 1. / 3.);
    pw.MF2third = pw.MFthird * pw.MFthird;
    double alpha_offset, epsilon_offset;
    euler_angles(&alpha_offset, &epsilon_offset, &pw, &ac, &ec);
    double fcut = .2 / M;
    for (int j = 0; j < n; j++) {
        double fj = f[j];
        if (fj > fcut) {
            amp_vec[j] = 0;
            phase_vec[j] = 0;
            hpf[2 * j] = hpf[2 * j + 1] = hcf[2 * j] = hcf[2 * j + 1] = 0;
            continue;
        }
        gwo_precalc_powers_ins(fj, M, &pw);
        double amp = (A0 * gwo_build_amp(fj, s, &st, &pw));
        double phase = gwo_build_phase(fj, s, &st, &pw);
        double d2[5], dm2[5];
        wigner_d(d2, dm2, &pw, s);
        double alpha, epsilon;
        euler_angles(&alpha, &epsilon, &pw, &ac, &ec);
        alpha = alpha + s->alpha0 - alpha_offset;
        epsilon = epsilon - epsilon_offset;
        twistup(alpha, d2, dm2, y, &hpf[2 * j], &hcf[2 * j]);
        phase = phase + 2. * epsilon;
        amp_vec[j] = amp / 2.;
        phase_vec[j] = phase;
    }
    double t_corr = s->shift_time ? time_shift(s, &st, &pw) : 0;
    for (int j = 0; j < n; j++) {
        double arg = phase_vec[j] - tc * (f[j] - f_ref) - phic + 2 * M_PI * t_corr * f[j];
        double er = cos(arg), ei = -sin(arg);
        double ar = amp_vec[j] * hpf[2 * j], ai = amp_vec[j] * hpf[2 * j + 1];
        gwo_cmul(ar, ai, er, ei, &hp[2 * j], &hp[2 * j + 1]);
        ar = amp_vec[j] * hcf[2 * j];
        ai = amp_vec[j] * hcf[2 * j + 1];
        gwo_cmul(ar, ai, er, ei, &hc[2 * j], &hc[2 * j + 1]);
    }
    free(amp_vec);
    free(phase_vec);
    free(hpf);
    free(hcf);
    return 1;
}

int gwo_pv2_setup_summary(const gwo_params *p, double *out)
{
    gwo_src s;
    gwo_populate_source(&s, p, GWO_IMRPHENOMPV2);
    gwo_phenomd_state st;
    int sst = s.shift_time, ssp = s.shift_phase;
    s.shift_time = 0;
    s.shift_phase = 0;
    gwo_phenomd_prepare(&s, &st);
    s.shift_time = sst;
    s.shift_phase = ssp;
    gwo_pows pw;
    gwo_precalc_powers_pi(&pw);
    gwo_precalc_powers_ins(s.f_ref, s.M, &pw);
    double t_corr = time_shift(&s, &st, &pw);
    double v[] = {s.chil, s.chip, s.SP, s.SL, s.thetaJN, s.alpha0, s.zeta_polariz, s.phi_aligned, t_corr,
                  s.fRD, s.fdamp, gwo_dphase_mr(s.fRD, &s, &st.lam)};
    memcpy(out, v, sizeof(v));
    return (int)(sizeof(v) / sizeof(double));
}

/* Pieces of the twist-up at one frequency, for pinning tests: out = [cos beta, d2[5] (d^2_{m,2}, m = -2..2),
 * dm2[5] (d^2_{m,-2}), -2Y_2m(thetaJN)[5], alpha, epsilon (with the f_ref offsets and alpha0 applied),
 * hp factor (re, im), hc factor (re, im), L2PN] */
int gwo_pv2_pieces(const gwo_params *p, double f, double *out)
{
    gwo_src s;
    gwo_populate_source(&s, p, GWO_IMRPHENOMPV2);
    s.q = s.mass1 / s.mass2;
    euler_coeffs ac, ec;
    euler_coeffs_calc(&ac, &ec, &s);
    gwo_pows pw;
    gwo_precalc_powers_pi(&pw);
    pw.MFthird = pow(s.M * s.f_ref, 1. / 3.);
    pw.MF2third = pw.MFthird * pw.MFthird;
    double a_off, e_off;
    euler_angles(&a_off, &e_off, &pw, &ac, &ec);
    gwo_precalc_powers_ins(f, s.M, &pw);
    double d2[5], dm2[5], y[5];
    wigner_d(d2, dm2, &pw, &s);
    ylm2(s.thetaJN, y);
    double alpha, eps;
    euler_angles(&alpha, &eps, &pw, &ac, &ec);
    alpha = alpha + s.alpha0 - a_off;
    eps = eps - e_off;
    double hp[2], hc[2];
    twistup(alpha, d2, dm2, y, hp, hc);
    double L = L2PN(s.eta, &pw);
    out[0] = 1. / sqrt(1.0 + (s.SP / (L + s.SL)) * (s.SP / (L + s.SL)));
    for (int i = 0; i < 5; i++) {
        out[1 + i] = d2[i];
        out[6 + i] = dm2[i];
        out[11 + i] = y[i];
    }
    out[16] = alpha;
    out[17] = eps;
    out[18] = hp[0];
    out[19] = hp[1];
    out[20] = hc[0];
    out[21] = hc[1];
    out[22] = L;
    return 23;
}

/* the waveform before the polarisation rotation by zeta (waveform_generator.cpp:257-266 applies it) */
int gwo_pv2_waveform_unrotated(const double *f, int n, double *hp, double *hc, const gwo_params *p)
{
    gwo_src s;
    gwo_populate_source(&s, p, GWO_IMRPHENOMPV2);
    return gwo_pv2_waveform(f, n, hp, hc, &s);
}

/* the aligned-spin phase the t_c correction splines (IMRPhenomD's build_phase with the precessing final spin),
 * at one frequency, for pinning tests */
double gwo_pv2_aligned_phase(const gwo_params *p, double f)
{
    gwo_src s;
    gwo_populate_source(&s, p, GWO_IMRPHENOMPV2);
    gwo_phenomd_state st;
    s.shift_time = 0;
    s.shift_phase = 0;
    gwo_phenomd_prepare(&s, &st);
    gwo_pows pw;
    gwo_precalc_powers_pi(&pw);
    if (f < s.f1_phase)
        gwo_precalc_powers_ins(f, s.M, &pw);
    else
        pw.MFsixth = pow(s.M * f, 1. / 6);
    return gwo_build_phase(f, &s, &st, &pw);
}
