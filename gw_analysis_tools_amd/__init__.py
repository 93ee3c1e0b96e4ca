"""MI355X engine for the gw_analysis_tools hot path.

The compute lives in libgwat_hip.so (HIP for gfx950) behind the C ABI of
include/gwat_hip.h; this package only binds it with ctypes.  There is no CPU
fallback: importing works anywhere, but every compute call raises if the
shared object or a HIP device is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgwat_hip.so")

GWH_OK = 1
GWH_ERR_INVALID, GWH_ERR_HIP, GWH_ERR_UNSUPPORTED, GWH_ERR_NOMEM, GWH_ERR_BUSY = -1, -2, -3, -4, -5
IMRPHENOMD = 0
IMRPHENOMPV2 = 1
PPE_IMRPHENOMD_INSPIRAL = 2
PPE_IMRPHENOMD_IMR = 3
PPE_MAX = 4
GIMRPHENOMD = 4
PPE_IMRPHENOMPV2_INSPIRAL = 5
PPE_IMRPHENOMPV2_IMR = 6
GIMRPHENOMPV2 = 7
# theories mapped onto the ppE models (ppe_theory; ppE_utilities.cpp:158-360) and the redshift cosmologies
(THEORY_NONE, THEORY_DCS, THEORY_EDGB, THEORY_EDGB_HO_LO, THEORY_EDGB_GHOV1, THEORY_EDGB_GHOV2, THEORY_EDGB_GHOV3,
 THEORY_EXTRADIMENSION, THEORY_BHEVAPORATION, THEORY_TVG, THEORY_DIPRAD, THEORY_NONCOMM, THEORY_PNSERIES,
 THEORY_PPEALT, THEORY_MODDISPERSION) = range(15)
COSMO_PLANCK15, COSMO_PLANCK13, COSMO_WMAP9, COSMO_WMAP7, COSMO_WMAP5 = range(5)
FP64 = 0
MIXED = 1
FISHER_GRAM_FP32 = 2  # Fisher only: FP32 operands on the FP32 matrix cores
CACHE_BY_CONTENT, CACHE_BY_POINTER = 0, 1  # how single calls find their cached event
SIMPSONS = 0
GAUSSLEG = 1
DETECTORS = {"Hanford": 0, "Livingston": 1, "Virgo": 2, "Kagra": 3, "Indigo": 4, "CE": 5, "ET1": 6, "ET2": 7,
             "ET3": 8}
BASIS_PHENOMD_SKYAVG = 0
BASIS_PHENOMD_FULL = 1
BASIS_PHENOMD_MCMC = 2
BASIS_PV2_FULL = 3
BASIS_PV2_MCMC = 4

# Every entry point of include/gwat_hip.h (checked by the CPU test suite).
EXPORTS = [
    "gwh_init", "gwh_last_error", "gwh_synchronize", "gwh_version",
    "gwh_fourier_waveform", "gwh_fourier_detector_response", "gwh_log_likelihood",
    "gwh_log_likelihood_internal", "gwh_mcmc_likelihood_extrinsic", "gwh_fisher_numerical",
    "gwh_event_create", "gwh_event_destroy", "gwh_event_bytes", "gwh_loglike_batch",
    "gwh_loglike_batch_device", "gwh_loglike_batch_device_noalloc", "gwh_set_bins_per_lane", "gwh_fisher_batch",
    "gwh_event_create_weighted", "gwh_gauleg", "gwh_set_single_call_cache", "gwh_host_data_register",
    "gwh_host_data_release",
    "gwh_single_call_stats",
    "gwh_time_loglike_batch_device", "gwh_device_alloc", "gwh_device_free", "gwh_memcpy_h2d",
    "gwh_memcpy_d2h", "gwh_work_counts", "gwh_create_coherent_GW_detection",
    "gwh_kernel_timing_start", "gwh_kernel_timing_every", "gwh_kernel_timing_stop", "gwh_device_copy",
    "gwh_fisher_derivatives", "gwh_fisher_numerical_model", "gwh_fisher_derivatives_model",
    "gwh_unpack_parameters", "gwh_repack_parameters", "gwh_detector_id", "gwh_dtoa",
    "gwh_gps_to_gmst_radian", "gwh_loglike_multi_device", "gwh_autotune_bins_per_lane", "gwh_autotune_tiles",
    "gwh_set_tile_lanes",
    "gwh_calculate_snr_internal", "gwh_data_snr", "gwh_snr_batch",
    "gwh_broker_create", "gwh_broker_destroy", "gwh_broker_submit_loglike", "gwh_broker_submit_fisher",
    "gwh_broker_wait", "gwh_broker_stats",
    "gwh_loglike_graph_create", "gwh_graph_launch", "gwh_graph_synchronize", "gwh_graph_destroy",
    "gwh_pt_create", "gwh_pt_run", "gwh_pt_synchronize", "gwh_pt_state", "gwh_pt_last_step", "gwh_pt_counters",
    "gwh_pt_destroy",
    "gwh_maximized_log_likelihood_aligned_internal", "gwh_maximized_log_likelihood_unaligned_internal",
    "gwh_maximized_coal_log_likelihood_internal", "gwh_maximized_loglike_batch",
]


class GWHError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """gwh_params: flat POD view of gen_params_base<double> (include/gwat/util.h:121-389)."""
    _fields_ = [("mass1", ctypes.c_double), ("mass2", ctypes.c_double),
                ("spin1", ctypes.c_double * 3), ("spin2", ctypes.c_double * 3),
                ("DL", ctypes.c_double), ("tc", ctypes.c_double), ("phiRef", ctypes.c_double),
                ("f_ref", ctypes.c_double), ("incl", ctypes.c_double), ("RA", ctypes.c_double),
                ("DEC", ctypes.c_double), ("psi", ctypes.c_double), ("gmst", ctypes.c_double),
                ("chip", ctypes.c_double), ("phip", ctypes.c_double),
                ("shift_time", ctypes.c_int32), ("shift_phase", ctypes.c_int32),
                ("sky_average", ctypes.c_int32), ("dep_postmerger", ctypes.c_int32),
                ("ppe_nmod", ctypes.c_int32), ("ppe_theory", ctypes.c_int32),
                ("ppe_b", ctypes.c_double * 4), ("ppe_beta", ctypes.c_double * 4),
                ("gimr_dphi", ctypes.c_double * 14), ("gimr_dsigma", ctypes.c_double * 3),
                ("gimr_dbeta", ctypes.c_double * 2), ("gimr_dalpha", ctypes.c_double * 4),
                ("gimr_nlist", ctypes.c_int32), ("cosmology", ctypes.c_int32), ("gimr_list", ctypes.c_int8 * 16)]


class PTConfig(ctypes.Structure):
    """gwh_pt_config (include/gwat_hip.h)."""
    _fields_ = [("n_temps", ctypes.c_int), ("n_chains", ctypes.c_int), ("model", ctypes.c_int),
                ("precision", ctypes.c_int), ("basis", ctypes.c_int), ("swap_every", ctypes.c_int),
                ("guard", ctypes.c_int), ("steps_per_graph", ctypes.c_int), ("tol", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


PARAMS_DTYPE = np.dtype([("mass1", "f8"), ("mass2", "f8"), ("spin1", "f8", 3), ("spin2", "f8", 3), ("DL", "f8"),
                         ("tc", "f8"), ("phiRef", "f8"), ("f_ref", "f8"), ("incl", "f8"), ("RA", "f8"),
                         ("DEC", "f8"), ("psi", "f8"), ("gmst", "f8"), ("chip", "f8"), ("phip", "f8"),
                         ("shift_time", "i4"), ("shift_phase", "i4"), ("sky_average", "i4"),
                         ("dep_postmerger", "i4"), ("ppe_nmod", "i4"), ("ppe_theory", "i4"), ("ppe_b", "f8", 4),
                         ("ppe_beta", "f8", 4), ("gimr_dphi", "f8", 14), ("gimr_dsigma", "f8", 3),
                         ("gimr_dbeta", "f8", 2), ("gimr_dalpha", "f8", 4), ("gimr_nlist", "i4"),
                         ("cosmology", "i4"), ("gimr_list", "i1", 16)])
assert PARAMS_DTYPE.itemsize == ctypes.sizeof(Params) == 448


def make_params(**kw):
    p = Params()
    p.chip = -1.0
    p.phip = -1.0
    p.shift_time = 1
    p.shift_phase = 1
    for k, v in kw.items():
        if k in ("spin1", "spin2", "ppe_b", "ppe_beta", "gimr_dphi", "gimr_dsigma", "gimr_dbeta", "gimr_dalpha",
                 "gimr_list"):
            getattr(p, k)[:len(v)] = list(v)
        else:
            setattr(p, k, v)
    return p


def params_array(n):
    a = np.zeros(n, dtype=PARAMS_DTYPE)
    a["chip"] = -1.0
    a["phip"] = -1.0
    a["shift_time"] = 1
    a["shift_phase"] = 1
    return a


_lib = None


def lib():
    """Load libgwat_hip.so; raises when it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GWHError("libgwat_hip.so is not built (python -m gw_analysis_tools_amd.build); "
                       "the engine has no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    D, I, V = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    pp = ctypes.POINTER(Params)
    dpp = ctypes.POINTER(dp)
    sig = {
        "gwh_init": ([I], I), "gwh_last_error": ([], ctypes.c_char_p), "gwh_synchronize": ([], I),
        "gwh_version": ([], ctypes.c_char_p),
        "gwh_fourier_waveform": ([dp, I, dp, dp, I, pp], I),
        "gwh_fourier_detector_response": ([dp, I, dp, I, I, pp], I),
        "gwh_log_likelihood": ([dp, dp, dp, ctypes.c_size_t, pp, I, I], D),
        "gwh_log_likelihood_internal": ([dp, dp, dp, dp, dp, I, I, I], D),
        "gwh_mcmc_likelihood_extrinsic": ([I, pp, I, ip, dpp, dpp, dpp, dpp, I, I, ip, I, D], D),
        "gwh_fisher_numerical": ([dp, I, I, I, I, dp, I, pp, I, dp, D], I),
        "gwh_event_create": ([dp, I, I, ip, dp, dp, ip], I),
        "gwh_event_create_weighted": ([dp, I, I, ip, dp, dp, I, dp, I, ip], I),
        "gwh_gauleg": ([D, D, dp, dp, I], I),
        "gwh_set_single_call_cache": ([I], I),
        "gwh_host_data_register": ([I, ip, dpp, dpp, dpp, dpp, ip], I),
        "gwh_host_data_release": ([I], I),
        "gwh_single_call_stats": ([dp], I),
        "gwh_event_destroy": ([I], I), "gwh_event_bytes": ([I], ctypes.c_size_t),
        "gwh_loglike_batch": ([ip, I, I, I, V, I, dp], I),
        "gwh_loglike_batch_device": ([V, I, I, I, V, I, V, V], I),
        "gwh_loglike_batch_device_noalloc": ([V, I, I, I, V, I, V, V], I),
        "gwh_set_bins_per_lane": ([I], I),
        "gwh_fisher_batch": ([I, I, I, I, V, I, I, D, dp, dp], I),
        "gwh_time_loglike_batch_device": ([I, I, I, V, I, V, I, dp, dp], I),
        "gwh_device_alloc": ([ctypes.c_size_t], V), "gwh_device_free": ([V], I),
        "gwh_memcpy_h2d": ([V, V, ctypes.c_size_t], I), "gwh_memcpy_d2h": ([V, V, ctypes.c_size_t], I),
        "gwh_work_counts": ([I, I, V, I, ip], I),
        "gwh_create_coherent_GW_detection": ([ip, I, dp, I, I, pp, dp], I),
        "gwh_kernel_timing_start": ([I], I), "gwh_kernel_timing_every": ([I], I),
        "gwh_kernel_timing_stop": ([dp, ip], I),
        "gwh_device_copy": ([V, V, ctypes.c_size_t], I),
        "gwh_fisher_derivatives": ([dp, I, I, I, I, I, pp, I, D, dp], I),
        "gwh_fisher_numerical_model": ([dp, I, I, I, I, I, dp, I, pp, I, dp, D], I),
        "gwh_fisher_derivatives_model": ([dp, I, I, I, I, I, I, pp, I, D, dp], I),
        "gwh_unpack_parameters": ([pp, I, dp, ip], I), "gwh_repack_parameters": ([dp, I, pp], I),
        "gwh_detector_id": ([ctypes.c_char_p], I), "gwh_dtoa": ([D, D, D, I, I], D),
        "gwh_gps_to_gmst_radian": ([D], D),
        "gwh_loglike_multi_device": ([ip, V, I, I, V, I, V, V], I),
        "gwh_autotune_bins_per_lane": ([ip, V, I, I, V, I, V, I, ip], I),
        "gwh_autotune_tiles": ([ip, V, I, I, V, I, V, I, ip, ip], I),
        "gwh_set_tile_lanes": ([I], I),
        "gwh_calculate_snr_internal": ([dp, dp, dp, I, I, dp, I], D),
        "gwh_data_snr": ([dp, I, dp, dp, dp], D),
        "gwh_snr_batch": ([ip, I, I, I, V, I, dp], I),
        "gwh_broker_create": ([ctypes.c_size_t, I, ip], I), "gwh_broker_destroy": ([I], I),
        "gwh_broker_submit_loglike": ([I, ip, I, I, I, V, I, dp, ctypes.POINTER(ctypes.c_longlong)], I),
        "gwh_broker_submit_fisher": ([I, I, I, I, I, V, I, I, D, dp, dp, ctypes.POINTER(ctypes.c_longlong)], I),
        "gwh_broker_wait": ([I, ctypes.c_longlong, D, dp], I),
        "gwh_broker_stats": ([I, dp], I),
        "gwh_loglike_graph_create": ([ip, V, I, I, I, V, I, V, ip], I),
        "gwh_graph_launch": ([I, V], I), "gwh_graph_synchronize": ([I], I), "gwh_graph_destroy": ([I], I),
        "gwh_pt_create": ([I, ctypes.POINTER(PTConfig), dp, dp, dp, dp, pp, dp, ip], I),
        "gwh_pt_run": ([I, I], I), "gwh_pt_synchronize": ([I], I),
        "gwh_pt_state": ([I, dp, dp, ctypes.POINTER(ctypes.c_longlong)], I),
        "gwh_pt_last_step": ([I, dp, dp, ip, ip], I),
        "gwh_pt_counters": ([I, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                             ctypes.POINTER(ctypes.c_longlong)], I),
        "gwh_pt_destroy": ([I], I),
        "gwh_maximized_log_likelihood_aligned_internal": ([dp, dp, dp, dp, ctypes.c_size_t], D),
        "gwh_maximized_log_likelihood_unaligned_internal": ([dp, dp, dp, dp, dp, ctypes.c_size_t], D),
        "gwh_maximized_coal_log_likelihood_internal": ([dp, dp, dp, dp, ctypes.c_size_t, dp, dp], D),
        "gwh_maximized_loglike_batch": ([I, I, V, I, dp], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def last_error():
    return lib().gwh_last_error().decode(errors="replace")


def check(rc):
    if rc != GWH_OK:
        raise GWHError("gwat_hip error %d: %s" % (rc, last_error()))
    return rc


from .engine import *  # noqa: E402,F401,F403
