"""FFT-maximised likelihoods on the GPU (rocFFT) against the CPU restatement (oracle/gwo_maxll.py):
single-call statistics (aligned, unaligned, coalescence) and the batched intrinsic network likelihood, to
1e-9 relative (the FP64 likelihood bar); the coalescence variant's time and phase at the same maximum."""
import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from oracle import gwo_maxll as M
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _c0(shift=0.37):
    """one detector, 20-1024 Hz at 1/4 Hz (4017 bins), injection shifted in time so the peak is off zero lag"""
    f = W.uniform_grid(20., 1024., 0.25)
    psd = W.analytic_psd(f)
    inj = O.make_params(**dict(W.INJECTION, tc=shift))
    h = O.detector_response(f, inj, 0)
    data = h + W.noise(f, psd, W.rng(0, 4))
    return f, psd, data


def test_single_call_statistics_match_restatement():
    f, psd, data = _c0()
    p = O.make_params(**dict(W.INJECTION, mass1=35.5, tc=0.))
    r = O.detector_response(f, p, 0)
    a = G.maximized_log_likelihood_aligned_internal(data, psd, f, r)
    ra = M.aligned(data, psd, f, r)
    assert abs(a - ra) <= TOL * abs(ra), (a, ra)
    ll, tc, ph = G.maximized_coal_log_likelihood_internal(data, psd, f, r)
    rl, rtc, rph = M.coal(data, psd, f, r)
    assert abs(ll - rl) <= TOL * abs(rl) and tc == rtc and abs(ph - rph) < 1e-9, (ll, rl, tc, rtc, ph, rph)
    # the recovered arrival time is the planted shift, modulo the transform's period
    T = 1. / (f[len(f) // 2] - f[len(f) // 2 - 1])
    assert min(abs(tc - 0.37), abs(T - tc - 0.37)) < 2. / T * T / len(f) + 1e-3
    hp, hc = O.fourier_waveform(f, O.make_params(**dict(W.INJECTION, spin1=[0.3, 0.1, 0.2])), O.IMRPHENOMPV2)
    u = G.maximized_log_likelihood_unaligned_internal(data, psd, f, hp, hc)
    ru = M.unaligned(data, psd, f, hp, hc)
    assert abs(u - ru) <= TOL * abs(ru), (u, ru)
    assert a > 0.4 * 0.5 * 4 * O.simpsons_sum(f[1] - f[0], np.abs(r) ** 2 / psd)  # a real peak, not noise


PV2 = (G.IMRPHENOMPV2, G.PPE_IMRPHENOMPV2_IMR, G.GIMRPHENOMPV2)


def _network(model):
    f = W.uniform_grid(20., 512., 0.125)
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    kw = dict(W.INJECTION)
    if model in PV2:
        kw.update(spin1=[0.3, 0.1, 0.2], spin2=[-0.1, 0.2, 0.1])
    inj = O.make_params(**kw)
    data = []
    g = W.rng(1, 5)
    for d in dets:
        h = O.detector_response(f, inj, d, O.IMRPHENOMPV2 if model in PV2 else O.IMRPHENOMD)
        dt = O.dtoa(inj.RA, inj.DEC, inj.gmst, dets[0], d)
        data.append(h * np.exp(-2j * np.pi * dt * f) + W.noise(f, psd, g))
    return f, psd, dets, data


@pytest.mark.parametrize("model", [G.IMRPHENOMD, G.IMRPHENOMPV2, G.PPE_IMRPHENOMD_IMR, G.PPE_IMRPHENOMPV2_IMR,
                                   G.GIMRPHENOMPV2])
def test_intrinsic_network_batch_matches_restatement(model):
    f, psd, dets, data = _network(model)
    ev = G.Event(f, data, [psd] * 3, dets)
    P = W.perturbed_params(W.injection_params(1)[0], 12, step=2)
    if model in PV2:
        P = W.precessing_params(12)
    if model in (G.PPE_IMRPHENOMD_IMR, G.PPE_IMRPHENOMPV2_IMR):
        P["ppe_nmod"] = 1
        P["ppe_b"][:, 0] = -7.
        P["ppe_beta"][:, 0] = np.linspace(-3e-5, 3e-5, len(P))
    if model == G.GIMRPHENOMPV2:
        P["gimr_dphi"][:, 6] = np.linspace(-0.02, 0.02, len(P))
    out = G.maximized_loglike_batch(ev, P, model)
    assert np.all(np.isfinite(out))
    for i in range(0, len(P), 3):
        rec = P[i]
        kw = {k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names}
        ref = M.intrinsic_network(f, data, [psd] * 3, O.make_params(**kw), model)
        assert abs(out[i] - ref) <= TOL * abs(ref), (i, out[i], ref)
    # an instance alone gives its batch value (to rounding: the batched transform may use another plan)
    alone = G.maximized_loglike_batch(ev, P[4:5], model)[0]
    assert abs(alone - out[4]) <= 1e-12 * abs(out[4])
    ev.close()


def test_intrinsic_statistic_ignores_extrinsic_parameters_for_aligned_spins():
    f, psd, dets, data = _network(G.IMRPHENOMD)
    ev = G.Event(f, data, [psd] * 3, dets)
    P = W.perturbed_params(W.injection_params(1)[0], 3, step=9)
    Q = P.copy()
    Q["DL"] *= 3.0
    Q["RA"] += 1.0
    Q["DEC"] *= -0.5
    Q["tc"] += 0.2
    a = G.maximized_loglike_batch(ev, P, G.IMRPHENOMD)
    b = G.maximized_loglike_batch(ev, Q, G.IMRPHENOMD)
    assert np.allclose(a, b, rtol=1e-11, atol=0)
    ev.close()


def test_maximised_batch_zero_data_and_unknown_event():
    f = W.uniform_grid(20., 100., 0.25)
    psd = W.analytic_psd(f)
    ev = G.Event(f, [np.zeros(len(f), complex)], [psd], [0])
    P = W.perturbed_params(W.injection_params(1)[0], 2, step=1)
    out = G.maximized_loglike_batch(ev, P, G.IMRPHENOMD)  # zero data: statistic 0
    assert np.all(out == 0)
    with pytest.raises(G.GWHError):
        G.maximized_loglike_batch(12345, P, G.IMRPHENOMD)
    ev.close()
