"""GPU parity of the HIP path against the CPU oracle on seeded inputs.

Tolerances (the FP64 bar): waveform and response <= 1e-10 relative per bin,
log-likelihood <= 1e-9 relative (relative to max(|logL|, 1)); the mixed path
|delta logL| <= 1e-3 against FP64.
"""
import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu

WF_TOL = 1e-10
LL_TOL = 1e-9

CASES = [
    dict(mass1=36, mass2=29, spin1=[0, 0, 0.3], spin2=[0, 0, -0.1]),
    dict(mass1=80, mass2=12, spin1=[0, 0, -0.8], spin2=[0, 0, 0.5]),
    dict(mass1=5.5, mass2=4.2, spin1=[0, 0, 0.9], spin2=[0, 0, 0.9]),
    dict(mass1=150, mass2=140, spin1=[0, 0, -0.95], spin2=[0, 0, 0.2]),
]
EXTR = dict(DL=410., RA=1.37, DEC=-1.26, psi=2.66, incl=2.9, f_ref=20., gmst=W.GMST, tc=0.01, phiRef=0.3)


def both(**kw):
    return G.make_params(**kw), O.make_params(**kw)


def wf_close(a, b, tol=WF_TOL):
    nz = np.abs(b) > 0
    assert np.all(a[~nz] == 0), "bins above the cutoff must be exactly zero"
    rel = np.abs(a[nz] - b[nz]) / np.abs(b[nz])
    assert rel.max() < tol, rel.max()


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("flags", [dict(shift_time=1, shift_phase=1), dict(shift_time=0, shift_phase=0),
                                   dict(dep_postmerger=1)])
def test_phenomd_waveform(case, flags):
    f = W.uniform_grid(20., 1024., 0.25)
    g, o = both(**case, **EXTR, **flags)
    hp, hc = G.fourier_waveform(f, g)
    op, oc = O.fourier_waveform(f, o)
    wf_close(hp, op)
    wf_close(hc, oc)


@pytest.mark.parametrize("case", CASES[:3])
def test_pv2_waveform(case):
    f = W.uniform_grid(20., 1024., 0.25)
    kw = dict(case, **EXTR)
    kw["spin1"] = [0.4, -0.2, case["spin1"][2] * 0.5]
    kw["spin2"] = [-0.1, 0.3, case["spin2"][2] * 0.5]
    g, o = both(**kw)
    hp, hc = G.fourier_waveform(f, g, G.IMRPHENOMPV2)
    op, oc = O.fourier_waveform(f, o, O.IMRPHENOMPV2)
    wf_close(hp, op)
    wf_close(hc, oc)
    # reduced (chip, phip) parametrisation
    g2, o2 = both(**dict(kw, chip=0.5, phip=1.1))
    hp, hc = G.fourier_waveform(f, g2, G.IMRPHENOMPV2)
    op, oc = O.fourier_waveform(f, o2, O.IMRPHENOMPV2)
    wf_close(hp, op)


@pytest.mark.parametrize("model", [G.IMRPHENOMD, G.IMRPHENOMPV2])
@pytest.mark.parametrize("det", [0, 1, 2, 6])
def test_detector_response(model, det):
    f = W.uniform_grid(20., 1024., 0.25)
    kw = dict(CASES[0], **EXTR)
    if model == G.IMRPHENOMPV2:
        kw["spin1"] = [0.3, 0.1, 0.2]
    g, o = both(**kw)
    wf_close(G.fourier_detector_response(f, g, det, model), O.detector_response(f, o, det, model))


def _c0_data():
    f = W.uniform_grid(20., 1024., 0.25)
    psd = W.analytic_psd(f)
    inj_g, inj_o = both(**W.INJECTION)
    h = O.detector_response(f, inj_o, 0)
    data = h + W.noise(f, psd, W.rng(0, 0))
    return f, psd, data


def test_log_likelihood_single_detector():
    f, psd, data = _c0_data()
    P = W.config0(32)["params"]
    for i in range(0, 32, 4):
        kw = {k: (list(P[i][k]) if P[i][k].shape else P[i][k].item()) for k in P.dtype.names}
        g, o = both(**kw)
        a = G.log_likelihood(data, psd, f, g, 0)
        b = O.log_likelihood(data, psd, f, o, 0)
        assert abs(a - b) <= LL_TOL * max(abs(b), 1), (i, a, b)


def _net_setup(n_bins_df=0.125, fmax=2048.):
    f = W.uniform_grid(20., fmax, n_bins_df)
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    _, inj = both(**W.INJECTION)
    data = []
    g = W.rng(1, 0)
    for d in dets:
        h = O.detector_response(f, inj, d)
        dt = O.dtoa(inj.RA, inj.DEC, inj.gmst, dets[0], d)
        data.append(h * np.exp(-2j * np.pi * dt * f) + W.noise(f, psd, g))
    return f, psd, dets, data


def _oracle_params(rec):
    return O.make_params(**{k: (list(rec[k]) if rec[k].shape else rec[k].item())
                            for k in rec.dtype.names})


@pytest.mark.parametrize("model", [G.IMRPHENOMD, G.IMRPHENOMPV2])
def test_network_batch_fp64_matches_oracle(model):
    f, psd, dets, data = _net_setup()
    ev = G.Event(f, data, [psd] * 3, dets)
    P = W.perturbed_params(W.injection_params(1)[0], 24, step=3)
    if model == G.IMRPHENOMPV2:
        P = W.precessing_params(24)
    ll = G.loglike_batch(ev, P, model, G.FP64)
    for i in range(0, 24, 3):
        ref = O.network_loglike(f, data, [psd] * 3, dets, _oracle_params(P[i]), model)
        assert abs(ll[i] - ref) <= LL_TOL * max(abs(ref), 1), (i, ll[i], ref)
    ev.close()


@pytest.mark.parametrize("model", [G.IMRPHENOMD, G.IMRPHENOMPV2])
def test_mixed_within_tolerance_of_fp64(model):
    f, psd, dets, data = _net_setup()
    ev = G.Event(f, data, [psd] * 3, dets)
    P = W.perturbed_params(W.injection_params(1)[0], 256, step=5)
    if model == G.IMRPHENOMPV2:
        P = W.precessing_params(256)
    a = G.loglike_batch(ev, P, model, G.FP64)
    b = G.loglike_batch(ev, P, model, G.MIXED)
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    d = np.abs(a - b)
    print("mixed vs fp64 model %d: max |dlogL| %.3e  median %.3e  max rel %.3e  max |logL| %.3e"
          % (model, d.max(), np.median(d), (d / np.abs(a)).max(), np.abs(a).max()))
    assert d.max() <= 1e-3, d.max()
    ev.close()


def test_batch_result_independent_of_batch_composition():
    f, psd, dets, data = _net_setup()
    ev = G.Event(f, data, [psd] * 3, dets)
    P = W.perturbed_params(W.injection_params(1)[0], 40, step=7)
    for prec in (G.FP64, G.MIXED):
        full = G.loglike_batch(ev, P, G.IMRPHENOMD, prec)
        alone = np.array([G.loglike_batch(ev, P[i:i + 1], G.IMRPHENOMD, prec)[0] for i in (0, 17, 39)])
        assert np.array_equal(alone, full[[0, 17, 39]])
        sub = G.loglike_batch(ev, P[10:30], G.IMRPHENOMD, prec)
        assert np.array_equal(sub, full[10:30])
    ev.close()


def test_mcmc_likelihood_extrinsic_T_eff():
    f, psd, dets, data = _net_setup(0.25, 1024.)
    p = W.injection_params(1)[0]
    g = G.make_params(**{k: (list(p[k]) if p[k].shape else p[k].item()) for k in p.dtype.names})
    o = _oracle_params(p)
    for T in (0.0, 4.0):
        a = G.mcmc_likelihood_extrinsic(f, data, [psd] * 3, dets, g, T_eff=T)
        b = O.mcmc_likelihood_extrinsic(f, data, [psd] * 3, dets, o, T_eff=T)
        assert abs(a - b) <= LL_TOL * max(abs(b), 1), (T, a, b)


def test_log_likelihood_internal_and_multi_event_batch():
    f, psd, data = _c0_data()
    _, o = both(**W.INJECTION)
    r = O.detector_response(f, o, 0)
    a = G.log_likelihood_internal(data, psd, f, r)
    b = O.loglike_internal(data, psd, f, r)
    assert abs(a - b) <= LL_TOL * abs(b)
    # two events of different length in one batch
    f2, psd2, dets2, data2 = _net_setup(0.25, 1024.)
    e1 = G.Event(f, [data], [psd], [0])
    e2 = G.Event(f2, data2, [psd2] * 3, dets2)
    P = W.perturbed_params(W.injection_params(1)[0], 6, step=9)
    evs = [e1, e2, e1, e2, e2, e1]
    mixed = G.loglike_batch(e1, P, G.IMRPHENOMD, G.FP64, events=evs)
    for i, e in enumerate(evs):
        alone = G.loglike_batch(e, P[i:i + 1], G.IMRPHENOMD, G.FP64)[0]
        assert mixed[i] == alone
    e1.close()
    e2.close()


def test_errors_are_reported_not_fatal():
    f = W.uniform_grid(20., 100., 1.)
    g = G.make_params(**dict(CASES[0], **EXTR))
    with pytest.raises(G.GWHError):
        G.fourier_detector_response(f, g, 99)
    with pytest.raises(G.GWHError):
        G.fourier_waveform(f, g, model=99)
    # eta > 1/4 style garbage propagates as NaN like the reference, without faulting
    bad = G.make_params(**dict(CASES[0], **EXTR))
    bad.mass2 = -3
    hp, hc = G.fourier_waveform(f, bad)
    assert np.isnan(hp).any()


def test_snr_entry_points():
    """calculate_snr_internal, data_snr and the batched network SNR against the oracle"""
    f = W.uniform_grid(20., 1024., 0.25)
    psd = W.analytic_psd(f)
    g, o = both(**dict(CASES[0], **EXTR))
    r = O.detector_response(f, o, 0)
    a = G.calculate_snr_internal(psd, r, f)
    b = O.snr(psd, r, f)
    assert abs(a - b) <= 1e-12 * b, (a, b)
    # data_snr: sqrt((d|r)) with d = 2 r -> sqrt(2) SNR
    assert abs(G.data_snr(f, 2 * r, r, psd) - np.sqrt(2) * b) <= 1e-12 * b
    # batched network SNR = sqrt(sum over detectors of the single-detector SNR^2)
    f2, psd2, dets, data = _net_setup()
    ev = G.Event(f2, data, [psd2] * 3, dets)
    P = W.perturbed_params(W.injection_params(1)[0], 16, step=4)
    s64 = G.snr_batch(ev, P, G.IMRPHENOMD, G.FP64)
    smx = G.snr_batch(ev, P, G.IMRPHENOMD, G.MIXED)
    for i in range(0, 16, 5):
        op = _oracle_params(P[i])
        ref = np.sqrt(sum(O.snr(psd2, O.detector_response(f2, op, d), f2) ** 2 for d in dets))
        assert abs(s64[i] - ref) <= 1e-10 * ref, (s64[i], ref)
    assert np.max(np.abs(smx / s64 - 1)) <= 1e-8
    ev.close()
