"""CPU restatement of the FFT-maximised likelihoods (TEST INFRASTRUCTURE: only tests/ import it).

Follows src/mcmc_gw.cpp:
  * maximized_coal_Log_Likelihood_internal (:493-556): maximised over t_c and phi_c,
  * maximized_Log_Likelihood_aligned_spin_internal (:595-654): over t_c, phi_c and the distance,
  * maximized_Log_Likelihood_unaligned_spin_internal (:660-800): over t_c, phi_c, the distance and the
    polarisation (the two-detector-free statistic of arXiv:1603.02444),
  * the MCMC "intrinsic" network likelihood that sums them over detectors (:2607-2700), with the extrinsic
    parameters fixed as there (theta = phi = psi = incl = 0, phiRef = 1, tc = 1, f_ref = 10 for IMRPhenomD
    and 20 for IMRPhenomPv2).
The FFT is numpy's (pocketfft) forward transform, the direction of the reference's FFTW_FORWARD plans
(util.cpp:968); the maxima do not depend on the direction.  Simpson sums are the C oracle's.
"""
import numpy as np

from . import IMRPHENOMD, IMRPHENOMPV2, PV2_FAMILY, PPE_IMRPHENOMD_INSPIRAL, PPE_IMRPHENOMD_IMR, fourier_waveform, make_params, \
    simpsons_sum


def aligned(data, psd, f, response):
    """maximized_Log_Likelihood_aligned_spin_internal: 0.5 max_k |FFT(4 d* h / S)_k|^2 df^2 / (h|h)"""
    df = f[1] - f[0]
    HH = 4. * simpsons_sum(df, np.real(response * np.conj(response)) / psd)
    out = np.fft.fft(4. * np.conj(data) * response / psd)
    mx = np.max(out.real * out.real + out.imag * out.imag) * df * df
    return .5 * mx / HH


def coal(data, psd, f, response):
    """maximized_coal_Log_Likelihood_internal: (logL, tc, phiRef) at the first maximum of |FFT|"""
    n = len(f)
    df = f[n // 2] - f[n // 2 - 1]
    HH = 4. * simpsons_sum(df, np.real(response * np.conj(response)) / psd)
    out = np.fft.fft(4. * np.conj(data) * response / psd)
    g = np.abs(out)
    k = int(np.argmax(g))
    max_val = g[k] * df
    return -0.5 * (HH - 2 * max_val), k / n * (1. / df), float(np.angle(out[k]))


def unaligned(data, psd, f, hplus, hcross):
    """maximized_Log_Likelihood_unaligned_spin_internal: 0.25 max_k lambda_k"""
    df = f[1] - f[0]
    hp_root = np.sqrt(4. * simpsons_sum(df, np.real(hplus * np.conj(hplus)) / psd))
    hc_root = np.sqrt(4. * simpsons_sum(df, np.real(hcross * np.conj(hcross)) / psd))
    hpn, hcn = hplus / hp_root, hcross / hc_root
    op = np.fft.fft(4. * np.conj(data) * hpn / psd)
    oc = np.fft.fft(4. * np.conj(data) * hcn / psd)
    rp, rc = df * op, df * oc
    rp2 = df * df * (op.real * op.real + op.imag * op.imag)
    rc2 = df * df * (oc.real * oc.real + oc.imag * oc.imag)
    gam = np.real(rp * np.conj(rc))
    ipc = 4. * simpsons_sum(df, np.real(hpn * np.conj(hcn)) / psd)
    lam = (rp2 + rc2 - 2 * gam * ipc + np.sqrt((rp2 - rc2) * (rp2 - rc2) + 4. * (ipc * rp2 - gam) *
                                                (ipc * rc2 - gam))) / (1. - ipc * ipc)
    return .25 * np.max(lam)


def intrinsic_params(p, model):
    """the extrinsic values the reference's intrinsic likelihood fixes before generating (mcmc_gw.cpp:2611-2617,
    :2660-2666)"""
    q = make_params()
    for name, _ in p._fields_:
        v = getattr(p, name)
        if hasattr(v, "__len__"):
            getattr(q, name)[:] = list(v)
        else:
            setattr(q, name, v)
    q.psi, q.phiRef, q.incl, q.tc = 0., 1., 0., 1.
    q.f_ref = 20. if model in PV2_FAMILY else 10.
    return q


def intrinsic_network(f, data, psd, p, model):
    """network sum of the maximised statistic: aligned (IMRPhenomD and its ppE / gIMR variants, on the plus
    polarisation, to which the horizon response at theta = phi = psi = 0 is proportional) or unaligned
    (IMRPhenomPv2 and its variants)"""
    q = intrinsic_params(p, model)
    hp, hc = fourier_waveform(f, q, model)
    total = 0.
    for d in range(len(data)):
        if model in PV2_FAMILY:
            total += unaligned(data[d], psd[d], f, hp, hc)
        else:
            total += aligned(data[d], psd[d], f, hp)
    return total
