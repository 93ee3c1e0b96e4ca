/*
 * gwo.h -- CPU oracle for the gw_analysis_tools hot path (TEST INFRASTRUCTURE).
 *
 * This library is a plain-C FP64 restatement of the reference's
 * IMRPhenomD / IMRPhenomPv2 waveform, detector projection, Simpson inner
 * product / log-likelihood and numerical Fisher matrix.  It keeps the
 * reference's loop structure and operation order (one instance per call,
 * per-call setup including the 1003-knot QNM spline build), so it doubles as
 * the "reference CPU path" that is timed beside the GPU.
 *
 * It is the checker, never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.
 *
 * Units follow the reference: masses in solar masses, distance in Mpc,
 * frequencies in Hz, complex arrays interleaved (re, im).
 *
 * Pinning (which check fixes which function) is listed in oracle/README.md.
 */
#ifndef GWO_H
#define GWO_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GWO_IMRPHENOMD = 0,
    GWO_IMRPHENOMPV2 = 1,
    GWO_PPE_IMRPHENOMD_INSPIRAL = 2,
    GWO_PPE_IMRPHENOMD_IMR = 3,
    GWO_GIMRPHENOMD = 4,
    GWO_PPE_IMRPHENOMPV2_INSPIRAL = 5,
    GWO_PPE_IMRPHENOMPV2_IMR = 6,
    GWO_GIMRPHENOMPV2 = 7
};
/* the precessing family (waveform_generator.cpp:227: local_method.find("IMRPhenomPv2")) */
/* theories mapped onto ppE terms (ppE_utilities.cpp:158-360); EdGB_HO maps as EdGB */
enum {
    GWO_THEORY_NONE = 0, GWO_THEORY_DCS, GWO_THEORY_EDGB, GWO_THEORY_EDGB_HO_LO, GWO_THEORY_EDGB_GHOV1,
    GWO_THEORY_EDGB_GHOV2, GWO_THEORY_EDGB_GHOV3, GWO_THEORY_EXTRADIMENSION, GWO_THEORY_BHEVAPORATION,
    GWO_THEORY_TVG, GWO_THEORY_DIPRAD, GWO_THEORY_NONCOMM, GWO_THEORY_PNSERIES, GWO_THEORY_PPEALT,
    GWO_THEORY_MODDISPERSION, GWO_THEORY_COUNT
};
double gwo_z_from_dl(double dl_mpc, int cosmo);
int gwo_md_index(double alpha);
double gwo_md_distance(double z, double alpha);

#define GWO_IS_PV2(m) \
    ((m) == GWO_IMRPHENOMPV2 || (m) == GWO_PPE_IMRPHENOMPV2_INSPIRAL || (m) == GWO_PPE_IMRPHENOMPV2_IMR || \
     (m) == GWO_GIMRPHENOMPV2)

/* Flat view of gen_params_base<double> (include/gwat/util.h:121-389): the
 * fields the hot path reads. */
typedef struct {
    double mass1, mass2;          /* M_sun (detector frame) */
    double spin1[3], spin2[3];
    double DL;                    /* Mpc */
    double tc;                    /* s */
    double phiRef;
    double f_ref;                 /* Hz */
    double incl;                  /* inclination angle */
    double RA, DEC, psi, gmst;    /* rad */
    double chip, phip;            /* Pv2 reduced spin parameters; chip=-1 => use spin vectors */
    int shift_time, shift_phase, sky_average;
    int dep_postmerger;           /* closed-form ringdown fits instead of the QNM spline (util.h:197) */
    int ppe_nmod, ppe_theory; /* theory id (GWO_THEORY_*), 0: plain ppE */        /* ppE terms: Nmod, bppe, betappe (util.h:249-255) */
    double ppe_b[4], ppe_beta[4];
    double gimr_dphi[14], gimr_dsigma[3], gimr_dbeta[2], gimr_dalpha[4]; /* gIMR deltas (0: unmodified) */
    /* the gIMR deltas in a Fisher basis, in the reference's list order (phii, sigmai, betai, alphai), as codes:
     * PN order i -> i + 4, sigma_i -> 12 + i, beta_i -> 15 + i, alpha_i -> 17 + i */
    int gimr_nlist, cosmology; /* cosmology: GWH_COSMO_* (0: PLANCK15) */
    signed char gimr_list[16];
} gwo_params;

void gwo_default_params(gwo_params *p);

/* --- waveform (waveform_generator.cpp:103-294) --- */
int gwo_fourier_waveform(const double *f, int n, double *hplus, double *hcross,
                         int model, const gwo_params *p);
/* construct_amplitude / construct_phase wrappers (waveform_generator.cpp:537-666, :815-920) */
int gwo_fourier_amplitude(const double *f, int n, double *amp, int model, const gwo_params *p);
int gwo_fourier_phase(const double *f, int n, double *phase_plus, double *phase_cross,
                      int model, const gwo_params *p);

/* --- detector (detector_util.cpp:676-790, :900-1180) --- */
int gwo_detector_id(const char *name);   /* -1 if unsupported */
int gwo_antenna(int det, double ra, double dec, double psi, double gmst, double *fplus, double *fcross);
double gwo_dtoa(double ra, double dec, double gmst, int det1, int det2);
double gwo_gps_to_gmst_radian(double gps);

/* fourier_detector_response, equatorial branch (waveform_util.cpp:911-985, :1069-1088) */
int gwo_detector_response(const double *f, int n, double *response, int det, int model,
                          const gwo_params *p);

/* --- statistics (mcmc_gw.cpp:570-591, :801-868, :2374-2489; util.h:841-856) --- */
double gwo_simpsons_sum(double dx, int n, const double *integrand);
double gwo_loglike_internal(const double *data, const double *psd, const double *f,
                            const double *weights, const double *response, int n,
                            int log10f, int gaussleg);
double gwo_log_likelihood(const double *data, const double *psd, const double *f, int n,
                          const gwo_params *p, int det, int model);
/* Network likelihood on one shared grid; T_eff replaces the reference's
 * pointer-difference T (tc_ref = T_eff - tc). */
double gwo_mcmc_likelihood_extrinsic(const gwo_params *p, int model, const double *f, int n,
                                     const double *const *data, const double *const *psd,
                                     const int *dets, int ndet, double T_eff);
/* Same network likelihood with the waveform tc used as given (no T_eff mapping). */
double gwo_network_loglike(const gwo_params *p, int model, const double *f, int n, const double *const *data,
                           const double *const *psd, const int *dets, int ndet);
double gwo_network_loglike_ex(const gwo_params *p, int model, const double *f, int n, const double *const *data,
                              const double *const *psd, const int *dets, int ndet, const double *weights, int log10f,
                              int gaussleg);
void gwo_fisher_elements(const double *f, int n, int dim, const double *deriv, const double *psd, int gaussleg,
                         const double *weights, int log10f, double *out);
/* Reference CPU path over a batch: one instance per call, OpenMP over instances
 * (mcmc_sampler.cpp:4583-4600 parallelises chains the same way). */
int gwo_network_loglike_batch(const gwo_params *ps, int n_inst, int model, const double *f, int n,
                              const double *const *data, const double *const *psd, const int *dets, int ndet,
                              int nthreads, double *out);
double gwo_snr(const double *psd, const double *response, const double *f, int n);

/* --- Fisher (fisher.cpp:81-557, :1754-2781) --- */
/* method: generation-method string as in the reference ("IMRPhenomD",
 * "MCMC_IMRPhenomD", "IMRPhenomPv2", "MCMC_IMRPhenomPv2"); sky_average in p
 * selects the sky-averaged basis.  out is dim*dim row major.  epsilon > 0: absolute step
 * (fisher.cpp:158); epsilon < 0: step |epsilon|*|v_i| per basis coordinate (|epsilon| at v_i = 0). */
int gwo_fisher_numerical(const double *f, int n, const char *method, int det, int ref_det,
                         double *out, int dim, const gwo_params *p, int order,
                         const double *psd, double epsilon);
int gwo_calculate_derivatives(double *deriv /* dim*2n */, const double *f, int n, int dim,
                              int det, int ref_det, const char *method, const gwo_params *p,
                              int order, double epsilon);
int gwo_unpack_parameters(double *vec, int *log_factors, const gwo_params *p,
                          const char *method, int dim);
int gwo_repack_parameters(const double *vec, gwo_params *out, const char *method, int dim,
                          const gwo_params *orig);
int gwo_lu_invert(const double *in, double *out, int dim);
int gwo_pn_phase_coefficients(const gwo_params *p, int model, double *out);

/* --- internals exposed for pinning tests --- */
/* natural cubic spline through (x,y), evaluated (deriv=0) or differentiated (deriv=1) at xq */
double gwo_spline_eval(const double *x, const double *y, int n, double xq, int deriv);
/* PhenomD setup summary: out[0..] = fRD, fdamp, f1, f3, f1_phase, f2_phase,
 * delta0..4, beta0, beta1, alpha0, alpha1, tc_shift, phic, A0, M, eta, chi_pn */
int gwo_phenomd_setup_summary(const gwo_params *p, int model, double *out);
/* PhenomD piecewise pieces at frequency f (no shifts): amp (A/(A0 M^7/6)), phase, damp, dphase */
int gwo_phenomd_pieces(const gwo_params *p, double f, double *out);
/* Pv2 setup summary: chil, chip, SP, SL, thetaJN, alpha0, zeta, phi_aligned, t_corr */
int gwo_pv2_setup_summary(const gwo_params *p, double *out);
int gwo_pv2_pieces(const gwo_params *p, double f, double *out);
double gwo_pv2_aligned_phase(const gwo_params *p, double f);
int gwo_pv2_waveform_unrotated(const double *f, int n, double *hp, double *hc, const gwo_params *p);
double gwo_fcut(const gwo_params *p);

#ifdef __cplusplus
}
#endif
#endif
