/* Internal declarations shared by the oracle's translation units (TEST INFRASTRUCTURE). */
#ifndef GWO_INTERNAL_H
#define GWO_INTERNAL_H
#include <math.h>
#include "gwo.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Constants of include/gwat/util.h:17-80 and IMRPhenomD.cpp:25 */
#define GWO_GAMMA_E 0.5772156649015328606065120900824024310421
#define GWO_C_SI 299792458.
#define GWO_MSOL_SEC 4.925491025543575903411922162094833998e-6
#define GWO_MPC_SI 3.085677581491367278913937957796471611e22
#define GWO_LOG_64 4.15888308336
#define GWO_DOUBLE_COMP_THRESH 1e-10
#define GWO_SQRT_6 2.44948974278317788
#define GWO_LOG10 2.302585092994045684017991454684364208

/* source_parameters<double> subset (util.h:426-665), dimensionful in seconds */
typedef struct {
    double mass1, mass2, M, q;
    double spin1z, spin2z, spin1x, spin2x, spin1y, spin2y;
    double chirpmass, eta, chi_s, chi_a, chi_eff, chi_pn, DL, delta_mass;
    double fRD, fdamp, f1, f3, f1_phase, f2_phase, tc, A0;
    int shift_phase, shift_time, sky_average, precessing, dep_postmerger;
    double chil, chip, phip, f_ref, phi_aligned, incl_angle, phiRef;
    double alpha0, thetaJN, zeta_polariz, SP, SL;
    int ppe;                      /* 0 GR, 1 ppE_IMRPhenomD_Inspiral, 2 ppE_IMRPhenomD_IMR */
    int Nmod;
    double bppe[4], betappe[4];
    int gimr;                     /* gIMRPhenomD: the deltas below apply */
    double dphi[14], dsigma[3], dbeta[2], dalpha[4];
} gwo_src;

/* lambda_parameters<double> (IMRPhenomD.h:16-26) */
typedef struct { double rho[4], v2, gamma[4], sigma[5], beta[5], alpha[7]; } gwo_lambda;

/* useful_powers<double> (util.h:396-419) */
typedef struct {
    double MFthird, MFsixth, MF7sixth, MF2third, MF4third, MF5third, MFsquare, MF7third,
        MF8third, MFcube, MFminus_5third, MF3fourth;
    double PIsquare, PIcube, PIthird, PI2third, PI4third, PI5third, PI7third, PIminus_5third;
} gwo_pows;

/* Everything construct_waveform derives once per call (IMRPhenomD.cpp:410-470). */
typedef struct {
    gwo_lambda lam;
    double deltas[5];
    double amp_coeff[7];
    double phase_coeff[12];
    double phic, f_ref, tc_total, A0M;
} gwo_phenomd_state;

double gwo_mpc_sec(void);
void gwo_populate_source(gwo_src *s, const gwo_params *p, int model);
int gwo_theory_terms(const gwo_params *p, double M, double mc, double eta, double *bppe, double *betappe);
void gwo_phenomd_prepare(gwo_src *s, gwo_phenomd_state *st);
void gwo_precalc_powers_pi(gwo_pows *pw);
void gwo_precalc_powers_ins(double f, double M, gwo_pows *pw);
double gwo_build_amp(double f, gwo_src *s, gwo_phenomd_state *st, gwo_pows *pw);
double gwo_build_phase(double f, gwo_src *s, gwo_phenomd_state *st, gwo_pows *pw);
double gwo_phase_ins(double f, gwo_src *s, double *pn, gwo_lambda *lam, gwo_pows *pw);
double gwo_phase_int(double f, gwo_src *s, gwo_lambda *lam);
double gwo_phase_mr(double f, gwo_src *s, gwo_lambda *lam);
double gwo_amp_ins(double f, gwo_src *s, double *pn, gwo_lambda *lam, gwo_pows *pw);
double gwo_amp_int(double f, gwo_src *s, double *deltas);
double gwo_amp_mr(double f, gwo_src *s, gwo_lambda *lam);
double gwo_dphase_ins(double f, gwo_src *s, double *pn, gwo_lambda *lam);
double gwo_dphase_int(double f, gwo_src *s, gwo_lambda *lam);
double gwo_dphase_mr(double f, gwo_src *s, gwo_lambda *lam);
double gwo_damp_ins(double f, gwo_src *s, double *pn, gwo_lambda *lam);
double gwo_damp_mr(double f, gwo_src *s, gwo_lambda *lam);
double gwo_lambda_element(const gwo_src *s, int i);
double gwo_calculate_eta(double m1, double m2);
double gwo_calculate_chirpmass(double m1, double m2);
double gwo_calculate_mass1(double mc, double eta);
double gwo_calculate_mass2(double mc, double eta);
double gwo_A0_from_DL(double mc, double dl, int sky_average);
double gwo_DL_from_A0(double mc, double a0, int sky_average);

/* Pv2 (IMRPhenomP.cpp) */
int gwo_pv2_waveform(const double *f, int n, double *hp, double *hc, gwo_src *s);
void gwo_pv2_param_transform(gwo_src *s, int reduced);

/* complex helpers on interleaved pairs */
static inline void gwo_cmul(double ar, double ai, double br, double bi, double *cr, double *ci)
{
    *cr = ar * br - ai * bi;
    *ci = ar * bi + ai * br;
}

#endif
