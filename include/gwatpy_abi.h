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

/* gwatpy_wrapping.h:178-222 -- gen_params_base_py.  Fields the engine does not model
 * (equatorial_orientation, horizon_coord, NSflag1/2, ppE and gIMR modifications) make every later
 * use of the handle fail with GWH_ERR_UNSUPPORTED. */
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

/* gwatpy_wrapping.h:115-119 (the reference calls repack_parameters with original_params = NULL) */
void repack_parameters_py(double *parameters, void *gen_param, char *generation_method, int dim);

/* gwatpy_wrapping.h:36 and :39 */
double gps_to_GMST_radian_py(double gps);
double DTOA_DETECTOR_py(double RA, double DEC, double GMST_rad, char *det1, char *det2);

#ifdef __cplusplus
}
#endif
#endif
