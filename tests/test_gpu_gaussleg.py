"""Gauss-Legendre and log10-f integration rules on the GPU (Log_Likelihood_internal's GAUSSLEG branch,
mcmc_gw.cpp:821-856; MCMC_likelihood_extrinsic with the caller's weights, :2374-2489; calculate_snr_internal,
waveform_util.cpp:479-510; calculate_fisher_elements, fisher.cpp:2704-2781) against the CPU oracle, whose rules
are pinned by closed-form integrals in tests/test_oracle_pins.py.  Bars: logL <= 1e-9 relative, SNR <= 1e-10,
Fisher <= 1e-6 of sqrt(F_ii F_jj)."""
import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu

LL_TOL = 1e-9
EXTR = dict(DL=410., RA=1.37, DEC=-1.26, psi=2.66, incl=2.9, f_ref=20., gmst=W.GMST, tc=0.01, phiRef=0.3)
SRC = dict(mass1=36, mass2=29, spin1=[0, 0, 0.3], spin2=[0, 0, -0.1])


def gl(x1, x2, n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x2 - x1) * x + 0.5 * (x2 + x1), 0.5 * (x2 - x1) * w


def grids():
    """(f, weights, log10F): Gauss-Legendre nodes on a linear and on a log10 frequency axis"""
    f, w = gl(20., 400., 800)
    u, wu = gl(np.log10(20.), np.log10(400.), 800)
    return [(f, w, False), (10 ** u, wu, True)]


@pytest.mark.parametrize("k", [0, 1])
def test_log_likelihood_internal_gaussleg(k):
    f, w, lg = grids()[k]
    psd = W.analytic_psd(f)
    g, o = G.make_params(**SRC, **EXTR), O.make_params(**SRC, **EXTR)
    r = O.detector_response(f, o, 0)
    d = O.detector_response(f, O.make_params(**dict(SRC, mass1=36.2), **EXTR), 0)
    ref = O.loglike_internal(d, psd, f, r, weights=w, log10F=lg, gaussleg=True)
    got = G.log_likelihood_internal(d, psd, f, r, weights=w, log10F=lg, method=G.GAUSSLEG)
    assert abs(got - ref) <= LL_TOL * max(abs(ref), 1), (got, ref)
    # the rule matters: Simpson weights on the same nodes give another value
    assert abs(G.log_likelihood_internal(d, psd, f, r) - ref) > 1e-3 * abs(ref)


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("model", [G.IMRPHENOMD, G.IMRPHENOMPV2])
def test_mcmc_likelihood_extrinsic_with_gauss_legendre_weights(k, model):
    f, w, lg = grids()[k]
    psd = W.analytic_psd(f)
    src = SRC if model == G.IMRPHENOMD else dict(SRC, spin1=[.3, .1, .2], spin2=[-.1, .2, .1])
    dets = [0, 1, 2]
    o_inj = O.make_params(**src, **EXTR)
    data = [O.detector_response(f, o_inj, d) * np.exp(-2j * np.pi * O.dtoa(o_inj.RA, o_inj.DEC, o_inj.gmst, 0, d) * f)
            for d in dets]
    kw = dict(src, **dict(EXTR, tc=0.01, DL=430.))
    kw["mass1"] = 36.3
    g, o = G.make_params(**kw), O.make_params(**kw)
    T = 0.0
    got = G.mcmc_likelihood_extrinsic(f, data, [psd] * 3, dets, g, model=model, T_eff=T, weights=w, log10F=lg,
                                      method=G.GAUSSLEG)
    o.tc = T - o.tc  # mcmc_gw.cpp:2466-2468
    ref = O.network_loglike(f, data, [psd] * 3, dets, o, model, weights=w, log10F=lg, gaussleg=True)
    assert abs(got - ref) <= LL_TOL * max(abs(ref), 1), (got, ref)


@pytest.mark.parametrize("k", [0, 1])
def test_calculate_snr_internal_rules(k):
    f, w, lg = grids()[k]
    psd = W.analytic_psd(f)
    o = O.make_params(**SRC, **EXTR)
    r = O.detector_response(f, o, 0)
    hh = -2 * O.loglike_internal(np.zeros_like(r), psd, f, r, weights=w, log10F=lg, gaussleg=True)
    got = G.calculate_snr_internal(psd, r, f, weights=w, log10F=lg, method=G.GAUSSLEG)
    assert abs(got / np.sqrt(hh) - 1) < 1e-10
    # Simpson with the log10 Jacobian (calculate_snr_internal applies it under either rule)
    fu = np.logspace(np.log10(20.), np.log10(400.), 801)
    ru = O.detector_response(fu, o, 0)
    ps = W.analytic_psd(fu)
    want = np.sqrt(O.fisher_elements(fu, ru[None, :], ps, log10F=True)[0, 0])
    got = G.calculate_snr_internal(ps, ru, fu, log10F=True)
    assert abs(got / want - 1) < 1e-10, (got, want)


@pytest.mark.parametrize("k", [0, 1])
def test_batched_loglike_on_a_gauss_legendre_event(k):
    """a resident event with the GAUSSLEG rule (gwh_event_create_weighted): every instance of a batch equals the
    oracle's network likelihood under the same weights; mixed within 1e-3 of FP64"""
    f, w, lg = grids()[k]
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    o_inj = O.make_params(**SRC, **EXTR)
    data = [O.detector_response(f, o_inj, d) * np.exp(-2j * np.pi * O.dtoa(o_inj.RA, o_inj.DEC, o_inj.gmst, 0, d) * f)
            for d in dets]
    ev = G.Event(f, data, [psd] * 3, dets, weights=w, log10F=lg)
    P = W.perturbed_params(W.injection_params(1)[0], 24, step=5)
    ll = G.loglike_batch(ev, P, G.IMRPHENOMD, G.FP64)
    for i in range(0, len(P), 3):
        rec = P[i]
        o = O.make_params(**{q: (list(rec[q]) if rec[q].shape else rec[q].item()) for q in rec.dtype.names})
        ref = O.network_loglike(f, data, [psd] * 3, dets, o, weights=w, log10F=lg, gaussleg=True)
        assert abs(ll[i] - ref) <= LL_TOL * max(abs(ref), 1), (i, ll[i], ref)
    mx = G.loglike_batch(ev, P, G.IMRPHENOMD, G.MIXED)
    assert np.max(np.abs(mx - ll)) <= 1e-3
    ev.close()


def test_fisher_batch_on_a_gauss_legendre_event():
    """the batched Fisher on a Gauss-Legendre log10 grid: the Gram uses the event's weights times f ln 10
    (calculate_fisher_elements' GAUSSLEG branch, fisher.cpp:2736-2745), summed over detectors with the first
    as the arrival-time reference; the oracle's derivatives and elements under the same rule"""
    f, w, lg = grids()[1]
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    zeros = [np.zeros(len(f), complex)] * 3
    ev = G.Event(f, zeros, [psd] * 3, dets, weights=w, log10F=True)
    P = W.perturbed_params(W.injection_params(1)[0], 2, step=11, config=3)
    eps = 1e-5
    F = G.fisher_batch(ev, P, G.IMRPHENOMD, G.BASIS_PHENOMD_MCMC, 11, precision=G.FP64, epsilon=eps)
    for i in range(len(P)):
        rec = P[i]
        o = O.make_params(**{q: (list(rec[q]) if rec[q].shape else rec[q].item()) for q in rec.dtype.names})
        ref = sum(O.fisher_elements(f, O.calculate_derivatives(f, "MCMC_IMRPhenomD", d, dets[0], 11, o, epsilon=eps),
                                    psd, weights=w, log10F=True, gaussleg=True) for d in dets)
        dg = np.sqrt(np.abs(np.diag(ref)))
        err = np.max(np.abs(F[i] - ref) / np.outer(dg, dg))
        assert err <= 1e-6, err
    ev.close()
