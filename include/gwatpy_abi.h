/*
 * gwatpy_abi.h -- the gwatpy C entry points of the hot path, exported by libgwat_hip.so with the names
 * and argument order of include/gwat/gwatpy_wrapping.h in the reference, so gwatpy's ctypes bindings
 * (gwatpy/gwatpy/waveform_generator.py:10-43, gwatpy/gwatpy/mcmc_routines.py) can load this library.
 *
 * The parameter handle is opaque to Python in both libraries; here it is the engine's own record (the
 * reference returns a heap gen_params_base<double>, gwatpy_wrapping.cpp:577-700).  Status codes follow
 * gwat_hip.h (1 = success); double-valued calls return NaN on failure, with gwh_last_error() set.
 */
#ifndef GWATPY_ABI_H
#define GWATPY_ABI_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* gwatpy_wrapping.h:178-222 -- gen_params_base_py.  The ppE terms (Nmod, bppe, betappe; at most GWH_PPE_MAX)
 * and the gIMR index lists travel with the handle: the ppE / theory and gIMR methods read them, the other methods
 * ignore them as prep_source_parameters does.  Fields the engine does not model (equatorial_orientation,
 * horizon_coord, NSflag1/2), more than GWH_PPE_MAX ppE terms, or ppE terms without betappe make every later use
 * of the handle fail with GWH_ERR_UNSUPPORTED. */
void *gen_params_base_py(double mass1, double mass2, double *spin1, double *spin2, double Luminosity_Distance,
                         double incl_angle, double RA, double DEC, double psi, double gmst, double tc, double phiRef,
                         double f_ref, double theta_l, double phi_l, double theta, double phi, char *cosmology,
                         bool equatorial_orientation, bool horizon_coord, bool NSflag1, bool NSflag2,
                         bool dep_postmerger, bool shift_time, bool shift_phase, bool sky_average, double LISA_alpha0,
                         double LISA_phi0, int Nmod_phi, int Nmod_sigma, int Nmod_beta, int Nmod_alpha, int *phii,
                         int *sigmai, int *betai, int *alphai, double *delta_phi, double *delta_sigma,
                         double *delta_beta, double *delta_alpha, int Nmod, double *bppe, double *betappe);

/* gwatpy_wrapping.h:223 */
void gen_params_base_py_destructor(void *p);

/* gwatpy_wrapping.h:169-176 (generation_method "IMRPhenomD" or "IMRPhenomPv2") */
int fourier_waveform_py(double *frequencies, int length, double *wf_plus_real, double *wf_plus_imaginary,
                        double *wf_cross_real, double *wf_cross_imaginary, char *generation_method, void *parameters);

/* gwatpy_wrapping.h:128-134 */
int fourier_detector_response_py(double *frequencies, int length, double *response_real, double *response_imaginary,
                                 char *detector, char *generation_method, void *parameters);

/* gwatpy_wrapping.h:58-71.  Flattened [detector][bin] arrays at i * data_length[i] + j.  Detector letters
 * are matched as the reference wrapper matches them (defaults Hanford, Livingston, Virgo; a suffix
 * equal to "H", "L", "V", "K" or "C" replaces the default).  The reference's time offset T is a pointer
 * difference (mcmc_gw.cpp:2466); here T = 0, so the waveform uses t_c = -tc. */
double MCMC_likelihood_extrinsic_py(bool save_waveform, void *parameters, char *generation_method, int *data_length,
                                    double *frequencies, double *dataREAL, double *dataIMAG, double *psd,
                                    double *weights, char *integration_method, bool log10F, char *detectors,
                                    int num_detectors);

/* gwatpy_wrapping.h:224-240: a handle holding the modification lists of a sampler's parameter vector (ppE
 * exponents, gIMR index lists; the NS flags make the likelihood below fail with GWH_ERR_UNSUPPORTED), and its
 * destructor (which here also frees the handle) */
void *MCMC_modification_struct_py(int ppE_Nmod, double *bppe, int gIMR_Nmod_phi, int *gIMR_phii, int gIMR_Nmod_sigma,
                                  int *gIMR_sigmai, int gIMR_Nmod_beta, int *gIMR_betai, int gIMR_Nmod_alpha,
                                  int *gIMR_alphai, bool NSflag1, bool NSflag2);
void MCMC_modification_struct_py_destructor(void *mod_struct);

/* gwatpy_wrapping.h:41-56: the network likelihood of a sampler's parameter vector in the "MCMC_" basis of the
 * method (dimension = GR block + modification parameters), as MCMC_prep_params (mcmc_gw.cpp:2492-2567) fills it:
 * f_ref = 20 Hz, shift_time and shift_phase on, gmst = 0 (the reference reads its module's mcmc_gmst, zero
 * unless a sampler run has set it), dCS / EdGB length scales in km raised to the theory's parameter; then as
 * MCMC_likelihood_extrinsic_py. */
double MCMC_likelihood_extrinsic_pyv2(bool save_waveform, double *parameters, void *mod_struct, int dimension,
                                      char *generation_method, int *data_length, double *frequencies,
                                      double *dataREAL, double *dataIMAG, double *psd, double *weights,
                                      char *integration_method, bool log10F, char *detectors, int num_detectors);

/* gwatpy_wrapping.h:115-119 (the reference calls repack_parameters with original_params = NULL) */
void repack_parameters_py(double *parameters, void *gen_param, char *generation_method, int dim);

/* gwatpy_wrapping.h:36 and :39 */
double gps_to_GMST_radian_py(double gps);
double DTOA_DETECTOR_py(double RA, double DEC, double GMST_rad, char *det1, char *det2);

#ifdef __cplusplus
}
#endif
#endif
