/*
 * gwat_hip.h -- C ABI of the MI355X engine for the gw_analysis_tools hot path:
 * IMRPhenomD / IMRPhenomPv2 frequency-domain waveforms, projection onto
 * terrestrial detectors, the Simpson-rule inner product behind the MCMC
 * log-likelihood, and the central-difference numerical Fisher matrix.
 *
 * Plain C types only (pointers, sizes, POD structs); every entry point is
 * thread-safe.  Status convention follows the reference: 1 (GWH_OK) means
 * success; negative values are errors described by gwh_last_error().
 * Functions returning a double return NaN on error.  NaN produced by the
 * physics (eta > 1/4, f_ref = 0 with shift_phase, PSD = 0, ...) propagates
 * into the result exactly as in the reference.
 *
 * Complex arrays are interleaved (re, im) doubles, matching
 * std::complex<double> storage.  Units follow the reference: masses in solar
 * masses (detector frame), distance in Mpc, times in seconds, angles in
 * radians, frequencies in Hz.
 */
#ifndef GWAT_HIP_H
#define GWAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GWH_OK 1
#define GWH_ERR_INVALID (-1)
#define GWH_ERR_HIP (-2)
#define GWH_ERR_UNSUPPORTED (-3)
#define GWH_ERR_NOMEM (-4)
#define GWH_ERR_BUSY (-5)

/* waveform models (the IMRPhenomD / IMRPhenomPv2 branches of
 * fourier_waveform, src/waveform_generator.cpp:129-266) */
#define GWH_IMRPHENOMD 0
#define GWH_IMRPHENOMPV2 1
/* IMRPhenomD with ppE phase terms sum_i betappe[i] (pi Mc f)^(bppe[i]/3) (src/ppE_IMRPhenomD.cpp:22-103):
 * in the inspiral phase only ("ppE_IMRPhenomD_Inspiral"; the C1 connections carry it into the later
 * regions) or in all three regions ("ppE_IMRPhenomD_IMR"). */
#define GWH_PPE_IMRPHENOMD_INSPIRAL 2
#define GWH_PPE_IMRPHENOMD_IMR 3
#define GWH_PPE_MAX 4 /* ppE terms per parameter set */
/* gIMRPhenomD (src/gIMRPhenomD.cpp): IMRPhenomD with the static PN phase coefficients, the sigma / beta /
 * alpha calibration fits and the negative-order inspiral phase terms scaled by (1 + delta) */
#define GWH_GIMRPHENOMD 4
#define GWH_GIMR_LIST_MAX 16            /* gIMR deltas in one Fisher basis */
#define GWH_GIMR_CODE_PHI(i) ((i) + 4)   /* PN order i = -4 .. 9 -> 0 .. 13 (gimr_dphi) */
#define GWH_GIMR_CODE_SIGMA(i) (12 + (i)) /* sigma_2 .. sigma_4 -> 14 .. 16 (gimr_dsigma) */
#define GWH_GIMR_CODE_BETA(i) (15 + (i))  /* beta_2, beta_3 -> 17, 18 (gimr_dbeta) */
#define GWH_GIMR_CODE_ALPHA(i) (17 + (i)) /* alpha_2 .. alpha_5 -> 19 .. 22 (gimr_dalpha) */
#define GWH_GIMR_NCODES 23
/* the same three modifications on IMRPhenomPv2 (src/ppE_IMRPhenomP.cpp, src/gIMRPhenomP.cpp): the aligned-spin
 * phase of the precessing model carries the terms ("ppE_IMRPhenomPv2_Inspiral", "ppE_IMRPhenomPv2_IMR",
 * "gIMRPhenomPv2"; dispatch src/waveform_generator.cpp:227-246) */
#define GWH_PPE_IMRPHENOMPV2_INSPIRAL 5
#define GWH_PPE_IMRPHENOMPV2_IMR 6
#define GWH_GIMRPHENOMPV2 7

/* theories mapped onto the ppE models (src/ppE_utilities.cpp:84-135, :158-360); the model id is the ppE method the
 * reference maps the theory's method string to (all of these the _Inspiral variant, PNSeries / ppEAlt either) */
#define GWH_THEORY_NONE 0
#define GWH_THEORY_DCS 1            /* dCS: b = -1, 16 pi alpha^2 / M_src^4 * dCS_phase_factor */
#define GWH_THEORY_EDGB 2           /* EdGB (and EdGB_HO, whose terms the reference's mapping overwrites with EdGB's) */
#define GWH_THEORY_EDGB_HO_LO 3     /* EdGB_HO_LO: b = -7, the 0PN higher-order beta */
#define GWH_THEORY_EDGB_GHOV1 4     /* EdGB_GHOv1..v3: b = -7 (EdGB) and -5 (generic 1PN, betappe[1]) */
#define GWH_THEORY_EDGB_GHOV2 5
#define GWH_THEORY_EDGB_GHOV3 6
#define GWH_THEORY_EXTRADIMENSION 7 /* b = -13, from the extra-dimension length squared (seconds^2) */
#define GWH_THEORY_BHEVAPORATION 8  /* b = -13, from the evaporation rate */
#define GWH_THEORY_TVG 9            /* b = -13, from G-dot */
#define GWH_THEORY_DIPRAD 10        /* b = -7, from delta E-dot */
#define GWH_THEORY_NONCOMM 11       /* b = -1, from Lambda^2 */
#define GWH_THEORY_PNSERIES 12      /* b as given; beta_0 (M/Mc)^(b_0/3), beta_0 beta_i (M/Mc)^(b_i/3) */
#define GWH_THEORY_PPEALT 13        /* b as given; beta_i (M/Mc)^(b_i/3) */
#define GWH_THEORY_MODDISPERSION 14 /* b = ppe_b[0] (alpha = b/3 + 1 in 0, 0.5, .., 4), IMR variant; from A_alpha */
#define GWH_THEORY_COUNT 15

/* cosmologies of the redshift fits (src/util.cpp:356-382, include/gwat/D_Z_Config.h) */
#define GWH_COSMO_PLANCK15 0
#define GWH_COSMO_PLANCK13 1
#define GWH_COSMO_WMAP9 2
#define GWH_COSMO_WMAP7 3
#define GWH_COSMO_WMAP5 4

/* arithmetic of the batched kernels */
#define GWH_FP64 0  /* everything in FP64 */
#define GWH_MIXED 1 /* FP64 except the FP32 operations that are faster on gfx950: the IMRPhenomD merger-ringdown
                       exponential and the IMRPhenomPv2 reciprocal square root's seed */

/* integration rules of Log_Likelihood_internal (mcmc_gw.cpp:801-868) */
#define GWH_SIMPSONS 0
#define GWH_GAUSSLEG 1

/* detector ids (the terrestrial names accepted by detector_util.cpp:1083-1158) */
#define GWH_HANFORD 0
#define GWH_LIVINGSTON 1
#define GWH_VIRGO 2
#define GWH_KAGRA 3
#define GWH_INDIGO 4
#define GWH_CE 5
#define GWH_ET1 6
#define GWH_ET2 7
#define GWH_ET3 8

/* Fisher parameter bases (fisher.cpp:1841-2161 unpack_parameters) */
#define GWH_BASIS_PHENOMD_SKYAVG 0   /* "IMRPhenomD", sky_average: (A0, phiRef, tc, Mc, eta, chi_s, chi_a), log on 0,3,4 */
#define GWH_BASIS_PHENOMD_FULL 1     /* "IMRPhenomD": (RA, DEC, psi, incl, phiRef, tc, DL, Mc, eta, chi1, chi2), log on 6,7 */
#define GWH_BASIS_PHENOMD_MCMC 2     /* "MCMC_IMRPhenomD": (RA, sinDEC, psi, cos incl, phiRef, tc, lnDL, lnMc, eta, chi1, chi2) */
#define GWH_BASIS_PV2_FULL 3         /* "IMRPhenomPv2": (..., chi1z, chi2z, chip, phip), 13 parameters */
#define GWH_BASIS_PV2_MCMC 4         /* "MCMC_IMRPhenomPv2": 15 parameters with spherical spins */

/* Flat POD view of gen_params_base<double> (include/gwat/util.h:121-389):
 * the fields the hot path reads. */
typedef struct gwh_params {
    double mass1, mass2;
    double spin1[3], spin2[3];
    double DL;
    double tc;
    double phiRef;
    double f_ref;
    double incl;
    double RA, DEC, psi;
    double gmst;
    double chip, phip; /* chip <= -1 + 1e-10: precessing spins are taken from spin1/spin2 vectors */
    int32_t shift_time, shift_phase, sky_average, dep_postmerger;
    /* ppE terms (gen_params_base: Nmod, bppe, betappe), read only by the ppE models. ppe_theory (GWH_THEORY_*,
     * 0: none) maps theory parameters held in ppe_beta[0 .. ppe_nmod) (and ppe_b for PNSeries / ppEAlt /
     * ModDispersion) to the ppE terms, as the reference does for its theory method strings (dCS_IMRPhenomD ->
     * ppE_IMRPhenomD_Inspiral, src/ppE_utilities.cpp:158-360, src/waveform_generator.cpp:1379-1404) */
    int32_t ppe_nmod, ppe_theory;
    double ppe_b[GWH_PPE_MAX];
    double ppe_beta[GWH_PPE_MAX];
    /* gIMR modifications (gen_params_base: phii / delta_phi, sigmai / delta_sigma, betai / delta_beta,
     * alphai / delta_alpha, as index -> delta tables; 0 leaves a coefficient as in IMRPhenomD), read only by
     * GWH_GIMRPHENOMD */
    double gimr_dphi[14];  /* PN order i = -4 .. 9 at [i + 4] (8, 9: the log terms of orders 5 and 6) */
    double gimr_dsigma[3]; /* sigma_2 .. sigma_4 */
    double gimr_dbeta[2];  /* beta_2, beta_3 */
    double gimr_dalpha[4]; /* alpha_2 .. alpha_5 */
    /* the gIMR deltas that are Fisher parameters, in the reference's basis order (the phii list, then sigmai,
     * betai, alphai: src/fisher.cpp:2128-2158), as codes into the tables above (GWH_GIMR_CODE_*); the first
     * entry of a repeated code is the one the waveform reads, as check_list_id finds it */
    int32_t gimr_nlist;
    int32_t cosmology; /* redshift from D_L for the theory maps (gen_params_base::cosmology): GWH_COSMO_*, 0 PLANCK15 */
    int8_t gimr_list[GWH_GIMR_LIST_MAX];
} gwh_params;

/* ---- runtime ---- */
int gwh_init(int device);              /* idempotent; called implicitly by every entry point */
const char *gwh_last_error(void);      /* message of this thread's last failure */
int gwh_synchronize(void);
const char *gwh_version(void);

/* ---- single-call drop-in entry points (host buffers in and out) ---- */

/* fourier_waveform(double*, int, waveform_polarizations<double>*, std::string, gen_params_base<double>*)
 * -- include/gwat/waveform_generator.h:43-49 */
int gwh_fourier_waveform(const double *frequencies, int length, double *hplus, double *hcross, int model,
                         const gwh_params *params);

/* fourier_detector_response(T*, int, std::complex<T>*, std::string det, std::string method,
 * gen_params_base<T>*, T* times=NULL) -- include/gwat/waveform_util.h:472-479 (terrestrial,
 * equatorial branch) */
int gwh_fourier_detector_response(const double *frequencies, int length, double *response, int detector,
                                  int model, const gwh_params *params);

/* create_coherent_GW_detection(std::string* detectors, int, T** frequencies, int* lengths, bool reuse_WF,
 * gen_params_base<T>*, std::string method, std::complex<T>** responses) -- include/gwat/waveform_util.h:297-315
 * (reuse_WF path, waveform_util.cpp:152-184): one waveform, projected on every detector and shifted by
 * the arrival-time difference to detectors[0].  responses is num_detectors*2*length (detector-major). */
int gwh_create_coherent_GW_detection(const int *detectors, int num_detectors, const double *frequencies, int length,
                                     int model, const gwh_params *params, double *responses);

/* ---- the single-call entry points below (the reference's signatures) ----
 * Their host arrays become resident events, cached (LRU of 64).  By default a call finds its event by a hash
 * of the arrays' contents (safe for any caller, ~0.1 ms per MB hashed, outside the engine lock).  Arrays registered with
 * gwh_host_data_register (the sampler's data, PSD, frequency and weight arrays, e.g. the module-level MCMC arrays
 * of mcmc_gw.cpp:1466-1474), or every array after gwh_set_single_call_cache(GWH_CACHE_BY_POINTER), are found by
 * their addresses instead: the caller keeps them unchanged while registered (or in pointer mode).  A thread that
 * repeats a call on the same arrays finds the event again without taking the engine lock.  Likelihood calls made
 * concurrently from several threads are coalesced into shared launches (the reference's samplers call them from
 * OpenMP and std::thread workers): the calling thread that finds a launch lane free (two exist) launches every
 * queued call of its model together and hands the results out.  Each result is the same as the call made alone. */
#define GWH_CACHE_BY_CONTENT 0
#define GWH_CACHE_BY_POINTER 1
int gwh_set_single_call_cache(int mode);
/* No counterpart in the reference (its arrays are globals of mcmc_gw.cpp).  weights may be NULL; data entries
 * may be NULL (arrays that are not likelihood data). */
int gwh_host_data_register(int num_detectors, const int *data_length, const double *const *frequencies,
                           const double *const *data, const double *const *psd, const double *const *weights,
                           int *handle);
int gwh_host_data_release(int handle);
/* No counterpart in the reference.  Counters of the single likelihood calls, 9 doubles in the layout of
 * gwh_broker_stats: [0] launches, [1] calls launched, [3] most calls in one launch, [6] calls queued now (others 0). */
int gwh_single_call_stats(double *stats);

/* Log_Likelihood(std::complex<double>*, double* psd, double* freqs, size_t, gen_params_base<double>*,
 * std::string det, std::string method) -- include/gwat/mcmc_gw.h:210-217 */
double gwh_log_likelihood(const double *data, const double *psd, const double *frequencies, size_t length,
                          const gwh_params *params, int detector, int model);

/* Log_Likelihood_internal(std::complex<double>*, double*, double*, double* weights,
 * std::complex<double>* response, int, bool log10F, std::string integration_method)
 * -- include/gwat/mcmc_gw.h:276-284 */
double gwh_log_likelihood_internal(const double *data, const double *psd, const double *frequencies,
                                   const double *weights, const double *response, int length, int log10F,
                                   int integration_method);

/* MCMC_likelihood_extrinsic(bool, gen_params_base<double>*, std::string, int* data_length,
 * double** frequencies, std::complex<double>** data, double** psd, double** weights,
 * std::string integration_method, bool log10F, std::string* detectors, int num_detectors)
 * -- include/gwat/mcmc_gw.h:473-485.  The reference derives T from a pointer difference
 * (mcmc_gw.cpp:2466); here T_eff is explicit and the waveform uses tc_ref = T_eff - tc.
 * All detectors must share one frequency grid (the reference's reuse_WF path). */
double gwh_mcmc_likelihood_extrinsic(int save_waveform, const gwh_params *params, int model,
                                     const int *data_length, const double *const *frequencies,
                                     const double *const *data, const double *const *psd,
                                     const double *const *weights, int integration_method, int log10F,
                                     const int *detectors, int num_detectors, double T_eff);

/* fisher_numerical(double*, int, string method, string det, string ref_det, double** out, int dim,
 * gen_params_base<double>*, int order, int*, int*, double* noise) -- include/gwat/fisher.h:35-48.
 * out is dim*dim row major; psd is the one-sided PSD on the same grid; epsilon > 0 is the
 * absolute step of the unpacked basis (the reference uses 1e-8).  epsilon < 0 selects a relative step
 * instead, |epsilon| * |theta_i| for each basis coordinate theta_i (|epsilon| where theta_i is 0);
 * this form has no counterpart in the reference.  epsilon 0 or non-finite: GWH_ERR_INVALID.  The same
 * convention holds for every Fisher entry point below. */
int gwh_fisher_numerical(const double *frequencies, int length, int basis, int detector, int reference_detector,
                         double *out, int dim, const gwh_params *params, int order, const double *psd,
                         double epsilon);

/* calculate_snr_internal(double *psd, std::complex<double> *waveform, double *frequencies, int length,
 * std::string integration_method, double *weights, bool log10_freq) -- src/waveform_util.cpp:479-510:
 * sqrt(4 Simpson(|h|^2 / S)) with the step f[1]-f[0] (or the Gauss-Legendre weights). */
double gwh_calculate_snr_internal(const double *psd, const double *waveform, const double *frequencies, int length,
                                  int integration_method, const double *weights, int log10F);

/* data_snr(double *frequencies, int length, std::complex<double> *data, std::complex<double> *response,
 * double *psd) -- src/waveform_util.cpp:108-127: sqrt((d|r)). */
double gwh_data_snr(const double *frequencies, int length, const double *data, const double *response,
                    const double *psd);

/* calculate_derivatives(double **derivatives, double *frequencies, int length, int dimension,
 * std::string detector, std::string reference_detector, std::string gen_method,
 * gen_params_base<double> *parameters, int order) -- src/fisher.cpp:160-557.
 * deriv: dim x length complex (interleaved re, im), parameter-major, with the log factors applied. */
int gwh_fisher_derivatives(const double *frequencies, int length, int basis, int detector, int reference_detector,
                           int dim, const gwh_params *params, int order, double epsilon, double *deriv);

/* The same two with the waveform model given (fisher_numerical / calculate_derivatives with a ppE or gIMR method
 * string, src/fisher.cpp:42-60, :2083-2088, :2128-2158, :2424-2429, :2473-2505): for the ppE ids the basis is
 * the GR basis of `basis` with betappe[0 .. params->ppe_nmod) appended, dim = basis dimension + ppe_nmod; for
 * the gIMR ids the deltas named by params->gimr_list[0 .. gimr_nlist) are appended, dim = basis dimension +
 * gimr_nlist; at most 16 in all.  Every point of a batch needs the same appended count. */
int gwh_fisher_numerical_model(const double *frequencies, int length, int model, int basis, int detector,
                               int reference_detector, double *out, int dim, const gwh_params *params, int order,
                               const double *psd, double epsilon);
int gwh_fisher_derivatives_model(const double *frequencies, int length, int model, int basis, int detector,
                                 int reference_detector, int dim, const gwh_params *params, int order, double epsilon,
                                 double *deriv);

/* unpack_parameters / repack_parameters (src/fisher.cpp:1841-2161, :2162-2511): the basis vector of a
 * parameter set and back.  log_factors (optional, dim ints) marks the parameters differentiated in log. */
int gwh_unpack_parameters(const gwh_params *params, int basis, double *v, int *log_factors);
int gwh_repack_parameters(const double *v, int basis, gwh_params *params);

/* Detector helpers: the name lookup of src/detector_util.cpp:573-636 (-1 if unknown), DTOA_DETECTOR
 * (src/detector_util.cpp:676-790; seconds, detector1 minus detector2) and gps_to_GMST_radian
 * (src/util.cpp:1797-1846). */
int gwh_detector_id(const char *name);
double gwh_dtoa(double RA, double DEC, double gmst, int detector1, int detector2);
double gwh_gps_to_gmst_radian(double gps);

/* ---- resident events and batches (the MCMC / serving path) ---- */

/* Upload one event: shared frequency grid, per-detector complex data and PSD (host buffers,
 * data is ndet*2*length, psd is ndet*length, detector-major).  Data stays resident on the GPU
 * until gwh_event_destroy.  Returns an event id >= 0 in *event. */
int gwh_event_create(const double *frequencies, int length, int ndet, const int *detectors, const double *data,
                     const double *psd, int *event);
int gwh_event_destroy(int event);
size_t gwh_event_bytes(int event);

/* Network log-likelihood (the MCMC_likelihood_extrinsic semantics, waveform tc used as given)
 * for n instances.  events[i] selects the resident event of instance i (NULL: all use events0).
 * Host buffers in and out. */
int gwh_loglike_batch(const int *events, int events0, int model, int precision, const gwh_params *params, int n,
                      double *logl);

/* gwh_event_create with the integration rule of Log_Likelihood_internal (mcmc_gw.cpp:801-868) and
 * MCMC_likelihood_extrinsic's weights argument: integration_method GWH_SIMPSONS (weights ignored) or
 * GWH_GAUSSLEG (weights: length doubles shared by every detector, e.g. from gwh_gauleg); log10F: the grid is
 * a log10 frequency axis, so Gauss-Legendre weights are multiplied by f ln 10 (the likelihood ignores it under
 * Simpson's rule, as the reference does).  Batched likelihoods, SNRs and Fisher matrices on the event use
 * its rule. */
int gwh_event_create_weighted(const double *frequencies, int length, int ndet, const int *detectors,
                              const double *data, const double *psd, int integration_method, const double *weights,
                              int log10F, int *event);

/* gauleg(double x1, double x2, double x[], double w[], int n) -- src/ortho_basis.cpp:14-46: Gauss-Legendre
 * nodes (ascending) and weights of the n-point rule on [x1, x2].  Host code; no GPU needed. */
int gwh_gauleg(double x1, double x2, double *x, double *w, int n);

/* Network optimal SNR sqrt(sum_d (h_d|h_d)) of n instances on resident events (the SNR convention of
 * calculate_snr_internal, src/waveform_util.cpp:479-510, summed over the event's detectors). */
int gwh_snr_batch(const int *events, int events0, int model, int precision, const gwh_params *params, int n,
                  double *snr);

/* Same with device buffers already resident in HBM (d_params: n gwh_params, d_logl: n doubles),
 * enqueued on the given HIP stream (NULL: the engine's stream) without a host synchronisation.
 * d_events may be NULL (all instances use events0). */
int gwh_loglike_batch_device(const int *d_events, int events0, int model, int precision, const void *d_params,
                             int n, double *d_logl, void *stream);

/* No counterpart in the reference.  gwh_loglike_batch_device for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same kernels and
 * results, with the instance records, tile counts and partial sums in a fixed device array of
 * GWH_NOALLOC_SCRATCH_BYTES that the library holds on the device until the process ends.  Calls no hipMalloc,
 * hipFree or allocating library routine.  gwh_init must have been called first (initialisation allocates the
 * model tables); a launch whose scratch exceeds the array fails with GWH_ERR_NOMEM before anything is enqueued
 * (4096 instances of a 16225-bin event need ~7 MB; a 2.6e5-bin event ~13 kB per instance).  Every call on the
 * device shares the one array: calls that overlap on the device (other streams, other threads' streams) race. */
#define GWH_NOALLOC_SCRATCH_BYTES (32u << 20)
int gwh_loglike_batch_device_noalloc(const int *d_events, int events0, int model, int precision,
                                     const void *d_params, int n, double *d_logl, void *stream);

/* Multi-event batch on device buffers with the instance -> event map also given on the host
 * (events: n ids, host; d_events: the same ids in device memory).  The host copy sizes one flattened
 * (instance, tile) work list per tiling class (bins per lane x lanes per tile), so events of very different
 * lengths share one launch without empty work items. */
int gwh_loglike_multi_device(const int *events, const int *d_events, int model, int precision, const void *d_params,
                             int n, double *d_logl, void *stream);

/* Autotuner: for each event-size bucket (floor(log2 N)) in the batch, time the batch with that bucket's
 * events at 1, 2, 4 and 8 bins per lane (each event keeping its lanes per tile) and keep the fastest (median
 * of reps); the choice is stored for later events of that size and written to table[32] (0 = bucket absent).
 * Results computed after the call may differ from earlier ones in the last bits (a different tiling sums in a
 * different order).  The project's own tuning step: the reference has no counterpart (its likelihood is one
 * serial loop per call, mcmc_gw.cpp:801-868). */
int gwh_autotune_bins_per_lane(const int *events, const int *d_events, int model, int precision, const void *d_params,
                               int n, double *d_logl, int reps, int *table);

/* Autotuner over the whole tiling: as gwh_autotune_bins_per_lane, over lanes per tile {64, 128, 256, 512, 1024}
 * (the kernels' workgroup size) x bins per lane {1, 2, 4, 8}, 20 candidates per bucket.  bpl_table[32] and
 * lanes_table[32] (optional) receive the choices (0 = bucket absent). */
int gwh_autotune_tiles(const int *events, const int *d_events, int model, int precision, const void *d_params, int n,
                       double *d_logl, int reps, int *bpl_table, int *lanes_table);

/* Tuning of the bins-per-lane of the likelihood kernels (1, 2, 4 or 8; 0 = automatic). */
int gwh_set_bins_per_lane(int bpl);

/* Tuning of the lanes per tile (workgroup size) of the likelihood kernels (64, 128, 256, 512 or 1024;
 * 0 = automatic: the autotuned choice of the event's size bucket, else 256).  A tile holds lanes x bins-per-lane
 * bins. */
int gwh_set_tile_lanes(int lanes);

/* Batched network Fisher: n points, order 2 central differences in the given basis (with the ppE betas or
 * listed gIMR deltas appended for those models, as gwh_fisher_numerical_model), per-detector projection
 * summed over the event's detectors.  out: n*dim*dim (host).  The stencil numerators are FP64 always;
 * precision GWH_FP64 and GWH_MIXED contract them on the FP64 matrix cores (the marginal sigmas of these
 * matrices, normalised condition numbers 1e8-1e10, need FP64 elements), GWH_FISHER_GRAM_FP32 feeds FP32
 * operands to the FP32 matrix cores with FP64 accumulation across tiles (elements to ~1e-6 of
 * sqrt(F_ii F_jj); not for sigmas).  sigma (optional, n*dim): marginal standard deviations from the FP64
 * inverse of each Fisher matrix. */
#define GWH_FISHER_GRAM_FP32 2
int gwh_fisher_batch(int event, int model, int basis, int precision, const gwh_params *params, int n, int dim,
                     double epsilon, double *out, double *sigma);

/* Device-resident timing helper for benchmarks: time `reps` launches of the batched likelihood on
 * device buffers with HIP events on the engine stream; returns milliseconds per launch and the
 * average duration of the dominant (per-bin) kernel in *kernel_ms. */
int gwh_time_loglike_batch_device(int events0, int model, int precision, const void *d_params, int n,
                                  double *d_logl, int reps, double *ms_per_launch, double *kernel_ms);

/* Bins of every instance by IMRPhenomD region on `event` (counts: n * GWH_NREGION_COUNTS ints): entry 3 a + p
 * counts the bins at or below 0.2/M in amplitude region a (0: f < 0.014/M, 1: up to the amplitude peak f3,
 * 2: above) and phase region p (0: f < 0.018/M, 1: up to f_RD / 2, 2: above); entry 9 the bins above 0.2/M.
 * Used to price the algorithmic work of a batch. */
#define GWH_NREGION_COUNTS 10
int gwh_work_counts(int event, int model, const gwh_params *params, int n, int *counts);

/* ---- serving broker (no counterpart in the reference: its samplers call the likelihood from many
 * threads at once, mcmc_sampler.cpp:4585-4600; the broker coalesces such calls) ----
 * Requests from any number of host threads are batched continuously onto the GPU: likelihood requests
 * of the same model and precision are concatenated (each instance keeps its event) up to
 * max_batch_instances, two streams keep one batch in flight while the next is formed; Fisher requests go to
 * two workers that merge queued requests of the same event, model, basis, precision and step.  Work is
 * admitted only if resident events + likelihood scratch + Fisher scratch fit in hbm_budget_bytes (Fisher
 * passes shrink to fit; what cannot fit on an idle device fails with GWH_ERR_NOMEM).  Results are bit-identical
 * to the same request run alone.  Parameters are copied at submission; outputs are written before the ticket
 * completes; the events a request reads are held from submission until its ticket is waited for (or the broker
 * destroyed): gwh_event_destroy returns GWH_ERR_BUSY meanwhile. */
int gwh_broker_create(size_t hbm_budget_bytes, int max_batch_instances, int *broker);
int gwh_broker_destroy(int broker); /* drains pending work first */
int gwh_broker_submit_loglike(int broker, const int *events, int event0, int model, int precision,
                              const gwh_params *params, int n, double *logl, long long *ticket);
int gwh_broker_submit_fisher(int broker, int event, int model, int basis, int precision, const gwh_params *params,
                             int n, int dim, double epsilon, double *out, double *sigma, long long *ticket);
/* waits for a ticket (timeout_ms < 0: no limit); returns the request's status (GWH_ERR_BUSY on timeout);
 * latency_us (optional): submission to completion */
int gwh_broker_wait(int broker, long long ticket, double timeout_ms, double *latency_us);
/* stats[9]: batches, instances, Fisher requests, most requests merged into one batch, admission waits,
 * peak device bytes (resident events + broker scratch), tickets not yet waited for, Fisher batches, most Fisher
 * requests merged into one Fisher batch */
int gwh_broker_stats(int broker, double *stats);

/* ---- likelihoods maximised over extrinsic parameters (FFT of the matched-filter integrand, rocFFT) ---- */

/* maximized_Log_Likelihood_aligned_spin_internal(std::complex<double>* data, double* psd, double* frequencies,
 * std::complex<double>* detector_response, size_t length, fftw_outline*) -- src/mcmc_gw.cpp:595-654:
 * 0.5 max_k |FFT(4 d* h / S)_k|^2 df^2 / (h|h), df = f[1] - f[0] (maximum over time, phase and distance) */
double gwh_maximized_log_likelihood_aligned_internal(const double *data, const double *psd, const double *frequencies,
                                                     const double *response, size_t length);

/* maximized_Log_Likelihood_unaligned_spin_internal(data, psd, frequencies, hplus, hcross, length, plan)
 * -- src/mcmc_gw.cpp:660-800: also maximised over the polarisation, 0.25 max_k lambda_k */
double gwh_maximized_log_likelihood_unaligned_internal(const double *data, const double *psd,
                                                       const double *frequencies, const double *hplus,
                                                       const double *hcross, size_t length);

/* maximized_coal_Log_Likelihood_internal(data, psd, frequencies, detector_response, length, plan, double* tc,
 * double* phiRef) -- src/mcmc_gw.cpp:493-556: -0.5 ((h|h) - 2 max_k |FFT_k| df), df = f[N/2] - f[N/2-1];
 * tc = k_max / N / df and phiRef = arg FFT_{k_max} at the first maximum */
double gwh_maximized_coal_log_likelihood_internal(const double *data, const double *psd, const double *frequencies,
                                                  const double *response, size_t length, double *tc, double *phiRef);

/* The intrinsic MCMC likelihood (src/mcmc_gw.cpp:2607-2700): for every parameter set, with the extrinsic
 * values fixed as there (psi = incl = 0, phiRef = 1, tc = 1, f_ref = 10 Hz, or 20 Hz for IMRPhenomPv2), the
 * maximised statistic of each detector of a resident Simpson-rule event, summed over its detectors: the
 * aligned statistic on the plus polarisation for the IMRPhenomD models (the reference's horizon response at
 * theta = phi = psi = 0 is proportional to it, and the statistic is invariant under that factor), the
 * unaligned one on (h+, hx) for IMRPhenomPv2.  FP64; host parameters in, host values out. */
int gwh_maximized_loglike_batch(int event, int model, const gwh_params *params, int n, double *out);

/* ---- sampler steps as replayable HIP graphs ---- */

/* Likelihood step of a sampler that makes its own proposals: the launch sequence of
 * gwh_loglike_batch_device (d_events, events NULL) or gwh_loglike_multi_device (both given) on fixed device
 * buffers, captured once; every gwh_graph_launch reads d_params and writes d_logl as they are when it runs
 * (the per-chain likelihood calls of the reference sampler, mcmc_sampler.cpp:4585-4600, in one replay).
 * The events stay held (gwh_event_destroy returns GWH_ERR_BUSY) until gwh_graph_destroy.  Launches of one
 * graph must be ordered (one stream): they share the graph's scratch. */
int gwh_loglike_graph_create(const int *events, const int *d_events, int events0, int model, int precision,
                             const void *d_params, int n, double *d_logl, int *graph);
int gwh_graph_launch(int graph, void *stream); /* NULL: the graph's own stream */
int gwh_graph_synchronize(int graph);
int gwh_graph_destroy(int graph);

/* Parallel tempering on the device: n_temps x n_chains chains (instance i = temp * n_chains + chain) on one
 * resident event, one whole step per graph replay with no host round trip:
 *   proposal   gaussian_step (src/mcmc_sampler_internals.cpp:364-409): one dimension chosen uniformly,
 *              moved by widths[i * dim + d] * N(0, 1)
 *   prior      a uniform box [lower, upper] in the sampling basis (outside: rejected, no likelihood)
 *   MH         mcmc_step (src/mcmc_sampler_internals.cpp:35-145): reject iff (ll' - ll) / T < log U
 *   swaps      after every swap_every-th step, chain_swap / single_chain_swap
 *              (src/mcmc_sampler_internals.cpp:1086-1180) along every ladder, pairs (k, k+1) in increasing k:
 *              reject iff exp((ll1 - ll2) / T2 - (ll1 - ll2) / T1) < U
 * The sampling basis is "MCMC_IMRPhenomD" (GWH_BASIS_PHENOMD_MCMC, 11 parameters) or "MCMC_IMRPhenomPv2"
 * (GWH_BASIS_PV2_MCMC, 15), mapped to parameter sets by repack_parameters (src/fisher.cpp:2162-2511) over
 * `tmpl`, which supplies every field the basis does not hold.  Random numbers: Philox4x32-10 keyed by the
 * seed, counter (instance, step, stream); a run is reproducible from (seed, step).  In GWH_MIXED precision
 * with guard = 1, decisions within 2 tol of their threshold (in log-likelihood, scaled by the temperatures)
 * are re-decided with FP64 likelihoods inside the same graph, so the run takes the FP64 run's decisions
 * while the mixed error stays within tol.  steps_per_graph steps are captured in one graph. */
typedef struct gwh_pt_config {
    int n_temps, n_chains; /* 1..64, 1..1024 */
    int model, precision, basis;
    int swap_every;        /* 0: never */
    int guard;
    int steps_per_graph;   /* >= 1 */
    double tol;            /* the mixed likelihood error bound of the guard band (1e-3) */
    uint64_t seed;
} gwh_pt_config;
/* temps: n_temps; widths: n x dim; lower, upper: dim; theta0: n x dim.  Computes the initial likelihoods. */
int gwh_pt_create(int event, const gwh_pt_config *cfg, const double *temps, const double *widths,
                  const double *lower, const double *upper, const gwh_params *tmpl, const double *theta0, int *pt);
int gwh_pt_run(int pt, int steps); /* enqueues `steps` steps (a multiple of steps_per_graph); returns at once */
int gwh_pt_synchronize(int pt);
/* after the enqueued steps: states (n x dim), their log-likelihoods, steps done (any pointer may be NULL) */
int gwh_pt_state(int pt, double *theta, double *ll, long long *step);
/* the last step: proposals (n x dim), their log-likelihoods (NaN outside the prior), accepted (0/1),
 * swapped pairs ((n_temps - 1) x n_chains, pair k of chain c at k * n_chains + c) */
int gwh_pt_last_step(int pt, double *prop, double *llp, int *accept, int *swapped);
/* accepted proposals per chain (n), accepted swaps per pair ((n_temps - 1) x n_chains), FP64 re-evaluations */
int gwh_pt_counters(int pt, long long *accepts, long long *swaps, long long *guard_evals);
int gwh_pt_destroy(int pt);

/* device memory helpers for callers without their own allocator */
void *gwh_device_alloc(size_t bytes);
int gwh_device_free(void *ptr);
int gwh_memcpy_h2d(void *dst, const void *src, size_t bytes);
int gwh_memcpy_d2h(void *dst, const void *src, size_t bytes);

/* Measurement helpers (no counterpart in the reference).  Between start and stop, every batched
 * likelihood launch records a HIP event pair around its per-bin kernel on the launch stream (at most
 * max_launches pairs); stop synchronises and returns the summed kernel time and the launch count.
 * gwh_kernel_timing_every(k) (after start) times only every k-th launch, so that the event markers
 * leave the timed launches' wall clock nearly untouched. */
int gwh_kernel_timing_start(int max_launches);
int gwh_kernel_timing_every(int k);
int gwh_kernel_timing_stop(double *total_ms, int *launches);
/* A plain 16-byte-vector copy kernel of `bytes` (device to device, synchronous), the reference point
 * for scaling hardware traffic counters. */
int gwh_device_copy(void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
