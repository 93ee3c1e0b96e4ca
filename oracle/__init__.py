"""ctypes binding to the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; it is the checker, never the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

IMRPHENOMD = 0
IMRPHENOMPV2 = 1
PPE_IMRPHENOMD_INSPIRAL = 2
PPE_IMRPHENOMD_IMR = 3
GIMRPHENOMD = 4
PPE_IMRPHENOMPV2_INSPIRAL = 5
PPE_IMRPHENOMPV2_IMR = 6
GIMRPHENOMPV2 = 7
PV2_FAMILY = (IMRPHENOMPV2, PPE_IMRPHENOMPV2_INSPIRAL, PPE_IMRPHENOMPV2_IMR, GIMRPHENOMPV2)
# theories mapped onto the ppE models (ppe_theory; ppE_utilities.cpp:158-360) and the redshift cosmologies
(THEORY_NONE, THEORY_DCS, THEORY_EDGB, THEORY_EDGB_HO_LO, THEORY_EDGB_GHOV1, THEORY_EDGB_GHOV2, THEORY_EDGB_GHOV3,
 THEORY_EXTRADIMENSION, THEORY_BHEVAPORATION, THEORY_TVG, THEORY_DIPRAD, THEORY_NONCOMM, THEORY_PNSERIES,
 THEORY_PPEALT, THEORY_MODDISPERSION) = range(15)
COSMO_PLANCK15, COSMO_PLANCK13, COSMO_WMAP9, COSMO_WMAP7, COSMO_WMAP5 = range(5)

DETECTORS = {"Hanford": 0, "Livingston": 1, "Virgo": 2, "Kagra": 3, "Indigo": 4,
             "CE": 5, "ET1": 6, "ET2": 7, "ET3": 8}


class Params(ctypes.Structure):
    """Mirror of gwo_params (oracle/gwo.h)."""
    _fields_ = [("mass1", ctypes.c_double), ("mass2", ctypes.c_double),
                ("spin1", ctypes.c_double * 3), ("spin2", ctypes.c_double * 3),
                ("DL", ctypes.c_double), ("tc", ctypes.c_double), ("phiRef", ctypes.c_double),
                ("f_ref", ctypes.c_double), ("incl", ctypes.c_double), ("RA", ctypes.c_double),
                ("DEC", ctypes.c_double), ("psi", ctypes.c_double), ("gmst", ctypes.c_double),
                ("chip", ctypes.c_double), ("phip", ctypes.c_double),
                ("shift_time", ctypes.c_int), ("shift_phase", ctypes.c_int),
                ("sky_average", ctypes.c_int), ("dep_postmerger", ctypes.c_int),
                ("ppe_nmod", ctypes.c_int), ("ppe_theory", ctypes.c_int),
                ("ppe_b", ctypes.c_double * 4), ("ppe_beta", ctypes.c_double * 4),
                ("gimr_dphi", ctypes.c_double * 14), ("gimr_dsigma", ctypes.c_double * 3),
                ("gimr_dbeta", ctypes.c_double * 2), ("gimr_dalpha", ctypes.c_double * 4),
                ("gimr_nlist", ctypes.c_int), ("cosmology", ctypes.c_int), ("gimr_list", ctypes.c_int8 * 16)]


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


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        D = ctypes.c_double
        P = ctypes.POINTER
        dp = P(D)
        L.gwo_fourier_waveform.argtypes = [dp, ctypes.c_int, dp, dp, ctypes.c_int, P(Params)]
        L.gwo_fourier_amplitude.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, P(Params)]
        L.gwo_fourier_phase.argtypes = [dp, ctypes.c_int, dp, dp, ctypes.c_int, P(Params)]
        L.gwo_antenna.argtypes = [ctypes.c_int, D, D, D, D, dp, dp]
        L.gwo_dtoa.argtypes = [D, D, D, ctypes.c_int, ctypes.c_int]
        L.gwo_dtoa.restype = D
        L.gwo_gps_to_gmst_radian.argtypes = [D]
        L.gwo_gps_to_gmst_radian.restype = D
        L.gwo_detector_response.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, P(Params)]
        L.gwo_simpsons_sum.argtypes = [D, ctypes.c_int, dp]
        L.gwo_simpsons_sum.restype = D
        L.gwo_loglike_internal.argtypes = [dp, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.gwo_loglike_internal.restype = D
        L.gwo_log_likelihood.argtypes = [dp, dp, dp, ctypes.c_int, P(Params), ctypes.c_int, ctypes.c_int]
        L.gwo_log_likelihood.restype = D
        L.gwo_mcmc_likelihood_extrinsic.argtypes = [P(Params), ctypes.c_int, dp, ctypes.c_int, P(dp), P(dp),
                                                    P(ctypes.c_int), ctypes.c_int, D]
        L.gwo_mcmc_likelihood_extrinsic.restype = D
        L.gwo_network_loglike.argtypes = [P(Params), ctypes.c_int, dp, ctypes.c_int, P(dp), P(dp),
                                          P(ctypes.c_int), ctypes.c_int]
        L.gwo_network_loglike.restype = D
        L.gwo_network_loglike_batch.argtypes = [P(Params), ctypes.c_int, ctypes.c_int, dp, ctypes.c_int,
                                                P(dp), P(dp), P(ctypes.c_int), ctypes.c_int, ctypes.c_int, dp]
        L.gwo_network_loglike_ex.argtypes = [P(Params), ctypes.c_int, dp, ctypes.c_int, P(dp), P(dp),
                                             P(ctypes.c_int), ctypes.c_int, dp, ctypes.c_int, ctypes.c_int]
        L.gwo_network_loglike_ex.restype = D
        L.gwo_fisher_elements.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, dp, ctypes.c_int, dp]
        L.gwo_snr.argtypes = [dp, dp, dp, ctypes.c_int]
        L.gwo_snr.restype = D
        L.gwo_fisher_numerical.argtypes = [dp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp,
                                           ctypes.c_int, P(Params), ctypes.c_int, dp, D]
        L.gwo_calculate_derivatives.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_char_p, P(Params), ctypes.c_int, D]
        L.gwo_unpack_parameters.argtypes = [dp, P(ctypes.c_int), P(Params), ctypes.c_char_p, ctypes.c_int]
        L.gwo_repack_parameters.argtypes = [dp, P(Params), ctypes.c_char_p, ctypes.c_int, P(Params)]
        L.gwo_lu_invert.argtypes = [dp, dp, ctypes.c_int]
        L.gwo_spline_eval.argtypes = [dp, dp, ctypes.c_int, D, ctypes.c_int]
        L.gwo_spline_eval.restype = D
        L.gwo_phenomd_setup_summary.argtypes = [P(Params), ctypes.c_int, dp]
        L.gwo_phenomd_pieces.argtypes = [P(Params), D, dp]
        L.gwo_pv2_setup_summary.argtypes = [P(Params), dp]
        L.gwo_pv2_pieces.argtypes = [P(Params), D, dp]
        L.gwo_pv2_aligned_phase.argtypes = [P(Params), D]
        L.gwo_pv2_aligned_phase.restype = D
        L.gwo_pv2_waveform_unrotated.argtypes = [dp, ctypes.c_int, dp, dp, P(Params)]
        L.gwo_pn_phase_coefficients.argtypes = [P(Params), ctypes.c_int, dp]
        L.gwo_fcut.argtypes = [P(Params)]
        L.gwo_fcut.restype = D
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def fourier_waveform(f, p, model=IMRPHENOMD):
    f = _c(f)
    n = len(f)
    hp = np.zeros(2 * n)
    hc = np.zeros(2 * n)
    lib().gwo_fourier_waveform(_dp(f), n, _dp(hp), _dp(hc), model, ctypes.byref(p))
    return hp[0::2] + 1j * hp[1::2], hc[0::2] + 1j * hc[1::2]


def fourier_amplitude(f, p):
    f = _c(f)
    a = np.zeros(len(f))
    lib().gwo_fourier_amplitude(_dp(f), len(f), _dp(a), IMRPHENOMD, ctypes.byref(p))
    return a


def fourier_phase(f, p):
    f = _c(f)
    a = np.zeros(len(f))
    b = np.zeros(len(f))
    lib().gwo_fourier_phase(_dp(f), len(f), _dp(a), _dp(b), IMRPHENOMD, ctypes.byref(p))
    return a


def antenna(det, ra, dec, psi, gmst):
    fp = ctypes.c_double()
    fx = ctypes.c_double()
    lib().gwo_antenna(det, ra, dec, psi, gmst, ctypes.byref(fp), ctypes.byref(fx))
    return fp.value, fx.value


def dtoa(ra, dec, gmst, d1, d2):
    return lib().gwo_dtoa(ra, dec, gmst, d1, d2)


def gps_to_gmst_radian(gps):
    return lib().gwo_gps_to_gmst_radian(gps)


def detector_response(f, p, det, model=IMRPHENOMD):
    f = _c(f)
    r = np.zeros(2 * len(f))
    lib().gwo_detector_response(_dp(f), len(f), _dp(r), det, model, ctypes.byref(p))
    return r[0::2] + 1j * r[1::2]


def _cplx(a):
    a = np.asarray(a, dtype=np.complex128)
    out = np.empty(2 * len(a))
    out[0::2] = a.real
    out[1::2] = a.imag
    return out


def simpsons_sum(dx, integrand):
    x = _c(integrand)
    return lib().gwo_simpsons_sum(dx, len(x), _dp(x))


def loglike_internal(data, psd, f, response, weights=None, log10F=False, gaussleg=False):
    """Log_Likelihood_internal (mcmc_gw.cpp:801-868): Simpson, or Gauss-Legendre weights (gaussleg)"""
    d = _cplx(data)
    r = _cplx(response)
    s = _c(psd)
    f = _c(f)
    w = _c(weights) if weights is not None else None
    return lib().gwo_loglike_internal(_dp(d), _dp(s), _dp(f), _dp(w) if w is not None else None, _dp(r), len(f),
                                      int(log10F), int(gaussleg))


def log_likelihood(data, psd, f, p, det, model=IMRPHENOMD):
    d = _cplx(data)
    s = _c(psd)
    f = _c(f)
    return lib().gwo_log_likelihood(_dp(d), _dp(s), _dp(f), len(f), ctypes.byref(p), det, model)


def _ptr_array(arrs):
    return (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_dp(a) for a in arrs])


def network_loglike(f, data, psd, dets, p, model=IMRPHENOMD, weights=None, log10F=False, gaussleg=False):
    f = _c(f)
    dd = [_cplx(x) for x in data]
    ss = [_c(x) for x in psd]
    di = (ctypes.c_int * len(dets))(*dets)
    if weights is None and not log10F and not gaussleg:
        return lib().gwo_network_loglike(ctypes.byref(p), model, _dp(f), len(f), _ptr_array(dd), _ptr_array(ss),
                                         di, len(dets))
    w = _c(weights) if weights is not None else None
    return lib().gwo_network_loglike_ex(ctypes.byref(p), model, _dp(f), len(f), _ptr_array(dd), _ptr_array(ss), di,
                                        len(dets), _dp(w) if w is not None else None, int(log10F), int(gaussleg))


def fisher_elements(f, deriv, psd, weights=None, log10F=False, gaussleg=False):
    """calculate_fisher_elements (fisher.cpp:2704-2781) of derivatives deriv[dim, n] (complex)"""
    f = _c(f)
    dim, n = deriv.shape
    d = np.empty(2 * dim * n)
    d[0::2] = deriv.real.ravel()
    d[1::2] = deriv.imag.ravel()
    s = _c(psd)
    w = _c(weights) if weights is not None else None
    out = np.zeros(dim * dim)
    lib().gwo_fisher_elements(_dp(f), n, dim, _dp(d), _dp(s), int(gaussleg), _dp(w) if w is not None else None,
                              int(log10F), _dp(out))
    return out.reshape(dim, dim)


def mcmc_likelihood_extrinsic(f, data, psd, dets, p, model=IMRPHENOMD, T_eff=0.0):
    f = _c(f)
    dd = [_cplx(x) for x in data]
    ss = [_c(x) for x in psd]
    di = (ctypes.c_int * len(dets))(*dets)
    return lib().gwo_mcmc_likelihood_extrinsic(ctypes.byref(p), model, _dp(f), len(f), _ptr_array(dd),
                                               _ptr_array(ss), di, len(dets), T_eff)


def network_loglike_batch(f, data, psd, dets, plist, model=IMRPHENOMD, nthreads=1):
    f = _c(f)
    dd = [_cplx(x) for x in data]
    ss = [_c(x) for x in psd]
    di = (ctypes.c_int * len(dets))(*dets)
    arr = (Params * len(plist))(*plist)
    out = np.zeros(len(plist))
    used = lib().gwo_network_loglike_batch(arr, len(plist), model, _dp(f), len(f), _ptr_array(dd),
                                           _ptr_array(ss), di, len(dets), nthreads, _dp(out))
    return out, used


def snr(psd, response, f):
    r = _cplx(response)
    s = _c(psd)
    f = _c(f)
    return lib().gwo_snr(_dp(s), _dp(r), _dp(f), len(f))


def fisher_numerical(f, psd, method, det, ref_det, dim, p, order=2, epsilon=1e-8):
    f = _c(f)
    s = _c(psd)
    out = np.zeros(dim * dim)
    ok = lib().gwo_fisher_numerical(_dp(f), len(f), method.encode(), det, ref_det, _dp(out), dim,
                                    ctypes.byref(p), order, _dp(s), epsilon)
    if not ok:
        raise ValueError("unsupported fisher configuration")
    return out.reshape(dim, dim)


def calculate_derivatives(f, method, det, ref_det, dim, p, order=2, epsilon=1e-8):
    f = _c(f)
    n = len(f)
    out = np.zeros(2 * n * dim)
    ok = lib().gwo_calculate_derivatives(_dp(out), _dp(f), n, dim, det, ref_det, method.encode(),
                                         ctypes.byref(p), order, epsilon)
    if not ok:
        raise ValueError("unsupported fisher configuration")
    out = out.reshape(dim, n, 2)
    return out[..., 0] + 1j * out[..., 1]


def unpack_parameters(p, method, dim):
    v = np.zeros(dim)
    lf = (ctypes.c_int * dim)()
    lib().gwo_unpack_parameters(_dp(v), lf, ctypes.byref(p), method.encode(), dim)
    return v, np.array(list(lf), dtype=bool)


def repack_parameters(v, method, dim, template):
    out = Params()
    ctypes.pointer(out)[0] = template
    v = _c(v)
    lib().gwo_repack_parameters(_dp(v), ctypes.byref(out), method.encode(), dim, ctypes.byref(template))
    return out


def lu_invert(m):
    m = _c(m)
    d = m.shape[0]
    out = np.zeros_like(m)
    lib().gwo_lu_invert(_dp(m), _dp(out), d)
    return out


def spline_eval(x, y, xq, deriv=0):
    x = _c(x)
    y = _c(y)
    return lib().gwo_spline_eval(_dp(x), _dp(y), len(x), xq, deriv)


def phenomd_setup_summary(p, model=IMRPHENOMD):
    out = np.zeros(32)
    n = lib().gwo_phenomd_setup_summary(ctypes.byref(p), model, _dp(out))
    keys = ["fRD", "fdamp", "f1", "f3", "f1_phase", "f2_phase", "delta0", "delta1", "delta2", "delta3",
            "delta4", "beta0", "beta1", "alpha0", "alpha1", "tc_shift", "phic", "A0", "M", "eta", "chi_pn"]
    return dict(zip(keys, out[:n]))


def phenomd_pieces(p, f):
    out = np.zeros(16)
    lib().gwo_phenomd_pieces(ctypes.byref(p), f, _dp(out))
    keys = ["amp_ins", "amp_int", "amp_mr", "phase_ins", "phase_int", "phase_mr", "damp_ins", "damp_mr",
            "dphase_ins", "dphase_int", "dphase_mr"]
    return dict(zip(keys, out[:11]))


def pv2_setup_summary(p):
    out = np.zeros(16)
    n = lib().gwo_pv2_setup_summary(ctypes.byref(p), _dp(out))
    keys = ["chil", "chip", "SP", "SL", "thetaJN", "alpha0", "zeta", "phi_aligned", "t_corr", "fRD", "fdamp",
            "dphase_mr_at_fRD"]
    return dict(zip(keys, out[:n]))


def pv2_pieces(p, f):
    """twist-up pieces at one frequency (gwo_pv2_pieces)"""
    out = np.zeros(23)
    lib().gwo_pv2_pieces(ctypes.byref(p), f, _dp(out))
    return dict(cos_beta=out[0], d2=out[1:6], dm2=out[6:11], ylm=out[11:16], alpha=out[16], epsilon=out[17],
                hp_factor=out[18] + 1j * out[19], hc_factor=out[20] + 1j * out[21], L2PN=out[22])


def pv2_aligned_phase(p, f):
    """the aligned-spin (IMRPhenomD) phase of a precessing parameter set at f, as the t_c correction splines it"""
    return lib().gwo_pv2_aligned_phase(ctypes.byref(p), f)


def pv2_waveform_unrotated(f, p):
    """IMRPhenomPv2 h+, hx before the rotation by the polarisation angle zeta"""
    f = _c(f)
    hp, hc = np.zeros(2 * len(f)), np.zeros(2 * len(f))
    lib().gwo_pv2_waveform_unrotated(_dp(f), len(f), _dp(hp), _dp(hc), ctypes.byref(p))
    return hp[0::2] + 1j * hp[1::2], hc[0::2] + 1j * hc[1::2]


def fcut(p):
    return lib().gwo_fcut(ctypes.byref(p))


def hanford_o1_fitted_asd(f):
    """Fitted Hanford O1 amplitude spectral density (detector_util.cpp:412-420)."""
    avec = [47.8466, -92.1896, 35.9273, -7.61447, 0.916742, -0.0588089, 0.00156345]
    x = np.log(f)
    return np.sqrt(.8464) * np.exp(avec[0] + avec[1] * x + avec[2] * x * x + avec[3] * x * x * x +
                                   avec[4] * x * x * x * x + avec[5] * x * x * x * x * x +
                                   avec[6] * x * x * x * x * x * x)


def pn_phase_coefficients(p, model=IMRPHENOMD):
    out = np.zeros(12)
    lib().gwo_pn_phase_coefficients(ctypes.byref(p), model, _dp(out))
    return out


def z_from_dl(dl_mpc, cosmology=COSMO_PLANCK15):
    """redshift of a luminosity distance in Mpc (Z_from_DL, util.cpp:356-382)"""
    f = lib().gwo_z_from_dl
    f.restype, f.argtypes = ctypes.c_double, [ctypes.c_double, ctypes.c_int]
    return f(dl_mpc, cosmology)


def theory_terms(p):
    """the ppE exponents and coefficients p.ppe_theory maps p's theory parameters to (assign_mapping,
    ppE_utilities.cpp:158-360), for the source as the waveform computes it; None if unmappable"""
    ms = 4.925491025543575903411922162094833998e-6
    L = lib()
    for name in ("gwo_calculate_chirpmass", "gwo_calculate_eta"):
        getattr(L, name).restype = ctypes.c_double
        getattr(L, name).argtypes = [ctypes.c_double, ctypes.c_double]
    m1, m2 = p.mass1 * ms, p.mass2 * ms
    b, beta = np.zeros(4), np.zeros(4)
    L.gwo_theory_terms.restype = ctypes.c_int
    L.gwo_theory_terms.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    n = L.gwo_theory_terms(ctypes.addressof(p), m1 + m2, L.gwo_calculate_chirpmass(m1, m2),
                           L.gwo_calculate_eta(m1, m2), _dp(b), _dp(beta))
    return None if n < 0 else (b[:n], beta[:n])
