"""gwatpy's ctypes view of the hot path, bound to libgwat_hip.so.

The argument lists below are the ones gwatpy declares for libgwat (gwatpy/gwatpy/util.py:8-53 for
gen_params_base_py, gwatpy/gwatpy/waveform_generator.py:21-43, gwatpy/gwatpy/mcmc_routines.py:83-99);
pointing gwatpy's config.LIB at libgwat_hip.so gives the same bindings.  `gen_params` mirrors
gwatpy.util.gen_params_base's keyword interface for the fields on the hot path.
"""
import ctypes

import numpy as np

from . import LIB_PATH

_D, _I, _B, _V, _S = ctypes.c_double, ctypes.c_int, ctypes.c_bool, ctypes.c_void_p, ctypes.c_char_p
_DP, _IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        L.gen_params_base_py.argtypes = ([_D, _D, _DP, _DP] + [_D] * 13 + [_S] + [_B] * 8 + [_D, _D] + [_I] * 4 +
                                         [_IP] * 4 + [_DP] * 4 + [_I, _DP, _DP])
        L.gen_params_base_py.restype = _V
        L.gen_params_base_py_destructor.argtypes = [_V]
        L.gen_params_base_py_destructor.restype = None
        L.fourier_waveform_py.argtypes = [_DP, _I, _DP, _DP, _DP, _DP, _S, _V]
        L.fourier_waveform_py.restype = _I
        L.fourier_detector_response_py.argtypes = [_DP, _I, _DP, _DP, _S, _S, _V]
        L.fourier_detector_response_py.restype = _I
        L.MCMC_likelihood_extrinsic_py.argtypes = [_B, _V, _S, _IP, _DP, _DP, _DP, _DP, _DP, _S, _B, _S, _I]
        L.MCMC_likelihood_extrinsic_py.restype = _D
        L.MCMC_modification_struct_py.argtypes = [_I, _DP, _I, _IP, _I, _IP, _I, _IP, _I, _IP, _B, _B]
        L.MCMC_modification_struct_py.restype = _V
        L.MCMC_modification_struct_py_destructor.argtypes = [_V]
        L.MCMC_modification_struct_py_destructor.restype = None
        L.MCMC_likelihood_extrinsic_pyv2.argtypes = [_B, _DP, _V, _I, _S, _IP, _DP, _DP, _DP, _DP, _DP, _S, _B, _S, _I]
        L.MCMC_likelihood_extrinsic_pyv2.restype = _D
        L.repack_parameters_py.argtypes = [_DP, _V, _S, _I]
        L.repack_parameters_py.restype = None
        L.gps_to_GMST_radian_py.argtypes = [_D]
        L.gps_to_GMST_radian_py.restype = _D
        L.DTOA_DETECTOR_py.argtypes = [_D, _D, _D, _S, _S]
        L.DTOA_DETECTOR_py.restype = _D
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


class gen_params:
    """Owning wrapper of a gen_params_base_py handle (keyword names as gwatpy.util.gen_params_base)."""

    def __init__(self, mass1=1., mass2=1., spin1=(0., 0., 0.), spin2=(0., 0., 0.), Luminosity_Distance=100.,
                 incl_angle=0., RA=0., DEC=0., psi=0., gmst=0., tc=0., phiRef=0., f_ref=0., theta_l=0., phi_l=0.,
                 theta=0., phi=0., cosmology="PLANCK15", equatorial_orientation=False, horizon_coord=False,
                 NSflag1=False, NSflag2=False, dep_postmerger=False, shift_time=True, shift_phase=True,
                 sky_average=False, LISA_alpha0=0., LISA_phi0=0., Nmod=0, bppe=None, betappe=None,
                 phii=None, delta_phi=None, sigmai=None, delta_sigma=None, betai=None, delta_beta=None,
                 alphai=None, delta_alpha=None):
        self._s1 = np.ascontiguousarray(spin1, dtype=np.float64)
        self._s2 = np.ascontiguousarray(spin2, dtype=np.float64)
        null_i, null_d = ctypes.cast(None, _IP), ctypes.cast(None, _DP)
        self._g = {}
        for key, ii, dd in (("phi", phii, delta_phi), ("sigma", sigmai, delta_sigma), ("beta", betai, delta_beta),
                            ("alpha", alphai, delta_alpha)):
            if ii is not None:
                self._g[key] = (np.ascontiguousarray(ii, dtype=np.int32), np.ascontiguousarray(dd, dtype=np.float64))

        def lst(key):
            if key not in self._g:
                return 0, null_i, null_d
            ii, dd = self._g[key]
            return len(ii), ii.ctypes.data_as(_IP), _dp(dd)
        nph, ph_i, ph_d = lst("phi")
        nsg, sg_i, sg_d = lst("sigma")
        nbt, bt_i, bt_d = lst("beta")
        nal, al_i, al_d = lst("alpha")
        self._b = None if bppe is None else np.ascontiguousarray(bppe, dtype=np.float64)
        self._be = None if betappe is None else np.ascontiguousarray(betappe, dtype=np.float64)
        self.handle = lib().gen_params_base_py(
            mass1, mass2, _dp(self._s1), _dp(self._s2), Luminosity_Distance, incl_angle, RA, DEC, psi, gmst, tc,
            phiRef, f_ref, theta_l, phi_l, theta, phi, cosmology.encode(), equatorial_orientation, horizon_coord,
            NSflag1, NSflag2, dep_postmerger, shift_time, shift_phase, sky_average, LISA_alpha0, LISA_phi0,
            nph, nsg, nbt, nal, ph_i, sg_i, bt_i, al_i, ph_d, sg_d, bt_d, al_d, Nmod,
            null_d if self._b is None else _dp(self._b), null_d if self._be is None else _dp(self._be))

    def close(self):
        if self.handle:
            lib().gen_params_base_py_destructor(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fourier_waveform(f, method, params):
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = len(f)
    out = [np.zeros(n) for _ in range(4)]
    st = lib().fourier_waveform_py(_dp(f), n, *[_dp(o) for o in out], method.encode(), params.handle)
    return st, out[0] + 1j * out[1], out[2] + 1j * out[3]


def fourier_detector_response(f, detector, method, params):
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = len(f)
    re, im = np.zeros(n), np.zeros(n)
    st = lib().fourier_detector_response_py(_dp(f), n, _dp(re), _dp(im), detector.encode(), method.encode(),
                                            params.handle)
    return st, re + 1j * im


def mcmc_likelihood_extrinsic(params, method, f, data, psd, detectors, weights=None, integration="SIMPSONS",
                              log10F=False, save_waveform=True):
    """data / psd: lists per detector on the shared grid f (flattened [detector][bin] as gwatpy passes them)"""
    D = len(data)
    n = len(f)
    lens = (ctypes.c_int * D)(*[n] * D)
    ff = np.ascontiguousarray(np.tile(f, D), dtype=np.float64)
    dr = np.ascontiguousarray(np.concatenate([np.real(x) for x in data]))
    di = np.ascontiguousarray(np.concatenate([np.imag(x) for x in data]))
    ps = np.ascontiguousarray(np.concatenate(psd), dtype=np.float64)
    w = np.ascontiguousarray(np.tile(weights, D), dtype=np.float64) if weights is not None else np.ones(n * D)
    return lib().MCMC_likelihood_extrinsic_py(save_waveform, params.handle, method.encode(), lens, _dp(ff), _dp(dr),
                                              _dp(di), _dp(ps), _dp(w), integration.encode(), log10F,
                                              detectors.encode(), D)


def repack_parameters(v, params, method, dim):
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib().repack_parameters_py(_dp(v), params.handle, method.encode(), dim)


class modification_struct:
    """gwatpy.mcmc_routines.MCMC_modification_struct_py: the modification lists of a sampler's parameter vector"""

    def __init__(self, ppE_Nmod=0, bppe=(), gIMR_phii=(), gIMR_sigmai=(), gIMR_betai=(), gIMR_alphai=(),
                 NSflag1=False, NSflag2=False):
        def arr(v, t):
            return (t * max(len(v), 1))(*v) if len(v) else None
        self.handle = lib().MCMC_modification_struct_py(
            ppE_Nmod, arr(list(bppe), ctypes.c_double), len(gIMR_phii), arr(list(gIMR_phii), ctypes.c_int),
            len(gIMR_sigmai), arr(list(gIMR_sigmai), ctypes.c_int), len(gIMR_betai), arr(list(gIMR_betai), ctypes.c_int),
            len(gIMR_alphai), arr(list(gIMR_alphai), ctypes.c_int), NSflag1, NSflag2)

    def close(self):
        if self.handle:
            lib().MCMC_modification_struct_py_destructor(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mcmc_likelihood_extrinsic_v2(parameters, mod, method, f, data, psd, detectors, weights=None,
                                 integration="SIMPSONS", log10F=False, save_waveform=True):
    """MCMC_likelihood_extrinsic_pyv2: a parameter vector in the method's MCMC basis (mcmc_routines.py:126-158)"""
    D = len(data)
    n = len(f)
    lens = (ctypes.c_int * D)(*[n] * D)
    ff = np.ascontiguousarray(np.tile(f, D), dtype=np.float64)
    dr = np.ascontiguousarray(np.concatenate([np.real(x) for x in data]))
    di = np.ascontiguousarray(np.concatenate([np.imag(x) for x in data]))
    ps = np.ascontiguousarray(np.concatenate(psd), dtype=np.float64)
    w = np.ascontiguousarray(np.tile(weights, D), dtype=np.float64) if weights is not None else np.ones(n * D)
    v = np.ascontiguousarray(parameters, dtype=np.float64)
    return lib().MCMC_likelihood_extrinsic_pyv2(save_waveform, _dp(v), mod.handle, len(v), method.encode(), lens,
                                                _dp(ff), _dp(dr), _dp(di), _dp(ps), _dp(w), integration.encode(),
                                                log10F, detectors.encode(), D)
