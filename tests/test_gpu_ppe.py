"""ppE phase terms on IMRPhenomD (ppE_IMRPhenomD_Inspiral / ppE_IMRPhenomD_IMR, src/ppE_IMRPhenomD.cpp) and the gIMR
coefficient modifications (gIMRPhenomD, src/gIMRPhenomD.cpp), and the same on IMRPhenomPv2 (src/ppE_IMRPhenomP.cpp,
src/gIMRPhenomP.cpp), on the GPU
against the CPU oracle: waveforms <= 1e-10 relative per bin, network log-likelihood <= 1e-9 relative (FP64),
mixed within 1e-3 of FP64, zero coefficients equal to IMRPhenomD bit for bit, and the refusals (more than
GWH_PPE_MAX terms, ppE Fisher bases)."""
import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from gw_analysis_tools_amd import gwatpy_compat as P
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu

WF_TOL = 1e-10
LL_TOL = 1e-9
EXTR = dict(DL=410., RA=1.37, DEC=-1.26, psi=2.66, incl=2.9, f_ref=20., gmst=W.GMST, tc=0.01, phiRef=0.3)
MODELS = [G.PPE_IMRPHENOMD_INSPIRAL, G.PPE_IMRPHENOMD_IMR]
TERMS = [dict(ppe_nmod=1, ppe_b=[-7.], ppe_beta=[3e-5]),
         dict(ppe_nmod=2, ppe_b=[-7., -3.], ppe_beta=[3e-5, -3e-3]),
         dict(ppe_nmod=3, ppe_b=[-5., -1., 1.5], ppe_beta=[-1e-3, 2e-2, 0.3])]
CASES = [dict(mass1=36, mass2=29, spin1=[0, 0, 0.3], spin2=[0, 0, -0.1]),
         dict(mass1=80, mass2=12, spin1=[0, 0, -0.8], spin2=[0, 0, 0.5]),
         dict(mass1=5.5, mass2=4.2, spin1=[0, 0, 0.9], spin2=[0, 0, 0.9])]


def both(**kw):
    return G.make_params(**kw), O.make_params(**kw)


def wf_close(a, b, tol=WF_TOL):
    nz = np.abs(b) > 0
    assert np.all(a[~nz] == 0)
    rel = np.abs(a[nz] - b[nz]) / np.abs(b[nz])
    assert rel.max() < tol, rel.max()


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("terms", TERMS)
def test_ppe_waveform_matches_oracle(model, case, terms):
    f = W.uniform_grid(20., 1024., 0.25)
    g, o = both(**case, **EXTR, **terms)
    hp, hc = G.fourier_waveform(f, g, model)
    op, oc = O.fourier_waveform(f, o, model)
    wf_close(hp, op)
    wf_close(hc, oc)
    # and it is not GR
    gp, _ = G.fourier_waveform(f, g, G.IMRPHENOMD)
    assert np.max(np.abs(np.angle(hp[np.abs(op) > 0] / gp[np.abs(op) > 0]))) > 1e-3


@pytest.mark.parametrize("model", MODELS)
def test_ppe_zero_terms_equal_gr_and_limits(model):
    f = W.uniform_grid(20., 1024., 0.25)
    kw = dict(CASES[0], **EXTR)
    gr, _ = G.fourier_waveform(f, G.make_params(**kw), G.IMRPHENOMD)
    z, _ = G.fourier_waveform(f, G.make_params(**kw, ppe_nmod=2, ppe_b=[-7., -3.], ppe_beta=[0., 0.]), model)
    assert np.array_equal(gr, z)
    neg, _ = G.fourier_waveform(f, G.make_params(**kw, ppe_nmod=-1, ppe_b=[-7.], ppe_beta=[1.]), model)
    assert np.array_equal(gr, neg)  # Nmod < 0: no terms, as the reference's loop
    # GR models ignore the ppE fields (prep_source_parameters copies them for the ppE methods only)
    ign, _ = G.fourier_waveform(f, G.make_params(**kw, **TERMS[1]), G.IMRPHENOMD)
    assert np.array_equal(gr, ign)
    # more terms than the record holds: NaN, never a silently truncated model
    over, _ = G.fourier_waveform(f, G.make_params(**kw, ppe_nmod=G.PPE_MAX + 1, ppe_b=[-7.] * 4,
                                                  ppe_beta=[1e-5] * 4), model)
    assert np.all(np.isnan(over[:100]))


def _net():
    f = W.uniform_grid(20., 512., 0.125)
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    inj_g, inj_o = both(**W.INJECTION, ppe_nmod=1, ppe_b=[-7.], ppe_beta=[2e-5])
    data = []
    g = W.rng(1, 3)
    for d in dets:
        h = O.detector_response(f, inj_o, d, O.PPE_IMRPHENOMD_INSPIRAL)
        dt = O.dtoa(inj_o.RA, inj_o.DEC, inj_o.gmst, dets[0], d)
        data.append(h * np.exp(-2j * np.pi * dt * f) + W.noise(f, psd, g))
    return f, psd, dets, data


@pytest.mark.parametrize("model", MODELS)
def test_ppe_network_loglike_batch(model):
    f, psd, dets, data = _net()
    ev = G.Event(f, data, [psd] * 3, dets)
    P_ = W.perturbed_params(W.injection_params(1)[0], 48, step=7)
    rng = np.random.default_rng(11)
    P_["ppe_nmod"] = rng.integers(0, 3, len(P_))
    P_["ppe_b"][:, 0], P_["ppe_b"][:, 1] = -7., -3.
    P_["ppe_beta"][:, 0] = rng.normal(0, 3e-5, len(P_))
    P_["ppe_beta"][:, 1] = rng.normal(0, 3e-3, len(P_))
    ll = G.loglike_batch(ev, P_, model, G.FP64)
    for i in range(0, len(P_), 6):
        rec = P_[i]
        kw = {k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names}
        ref = O.network_loglike(f, data, [psd] * 3, dets, O.make_params(**kw), model)
        assert abs(ll[i] - ref) <= LL_TOL * max(abs(ref), 1), (i, ll[i], ref)
    mx = G.loglike_batch(ev, P_, model, G.MIXED)
    assert np.max(np.abs(mx - ll)) <= 1e-3
    # each instance alone gives the same FP64 value (batch independence with mixed term counts)
    for i in (0, 5, 17):
        assert G.loglike_batch(ev, P_[i:i + 1], model, G.FP64)[0] == ll[i]
    ev.close()


def test_ppe_refusals_and_gwatpy():
    f = W.uniform_grid(20., 512., 0.25)
    psd = W.analytic_psd(f)
    ev = G.Event(f, [np.zeros(len(f), complex)], [psd], [0])
    P_ = W.perturbed_params(W.injection_params(1)[0], 2, step=1)
    P_["ppe_nmod"] = 1
    P_["ppe_b"][:, 0] = -7.
    # the ppE Fisher bases carry betappe as parameters: the GR dimension alone is refused
    with pytest.raises(G.GWHError) as e:
        G.fisher_batch(ev, P_, G.PPE_IMRPHENOMD_IMR, G.BASIS_PHENOMD_FULL, 11)
    assert "ppE" in str(e.value)
    # gIMR bases: more deltas than the record holds, unequal counts in a batch, a code outside the tables
    Q = P_.copy()
    Q["gimr_nlist"] = -1
    with pytest.raises(G.GWHError) as e:
        G.fisher_batch(ev, Q, G.GIMRPHENOMD, G.BASIS_PHENOMD_FULL, 11)
    assert "gIMR" in str(e.value)
    Q["gimr_nlist"] = [1, 2]
    with pytest.raises(G.GWHError):
        G.fisher_batch(ev, Q, G.GIMRPHENOMD, G.BASIS_PHENOMD_FULL, 12)
    Q["gimr_nlist"] = 1
    Q["gimr_list"][:, 0] = 23
    with pytest.raises(G.GWHError):
        G.fisher_batch(ev, Q, G.GIMRPHENOMD, G.BASIS_PHENOMD_FULL, 12)
    ev.close()
    # gwatpy handles carry Nmod / bppe / betappe to the ppE methods
    p = P.gen_params(mass1=36, mass2=29, spin1=(0, 0, .3), spin2=(0, 0, -.1), Luminosity_Distance=410.,
                     incl_angle=2.9, RA=1.37, DEC=-1.26, psi=2.66, gmst=W.GMST, tc=0.01, phiRef=0.3, f_ref=20.,
                     Nmod=2, bppe=[-7., -3.], betappe=[3e-5, -3e-3])
    st, hp, hc = P.fourier_waveform(f, "ppE_IMRPhenomD_IMR", p)
    assert st == G.GWH_OK
    o = O.make_params(**CASES[0], **EXTR, **TERMS[1])
    op, oc = O.fourier_waveform(f, o, O.PPE_IMRPHENOMD_IMR)
    wf_close(hp, op)
    p.close()


# ---- gIMR (gIMRPhenomD.cpp) --------------------------------------------------------------------------------
GIMR = [dict(gimr_dphi=[2e-4, 0, -1e-4, 0, 0, 0.01, 0.05, -0.03, 0.02, 0.1, -0.05, 0.04, 0.2, -0.1]),
        dict(gimr_dsigma=[0.1, -0.2, 0.3], gimr_dbeta=[0.05, -0.1], gimr_dalpha=[0.1, 0.2, -0.1, 0.05]),
        dict(gimr_dphi=[0] * 4 + [0, 0, 0.03, 0, 0, 0, 0, 0, 0, 0], gimr_dalpha=[0, 0, 0, 0.3])]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("mods", GIMR)
def test_gimr_waveform_matches_oracle(case, mods):
    f = W.uniform_grid(20., 1024., 0.25)
    g, o = both(**case, **EXTR, **mods)
    hp, hc = G.fourier_waveform(f, g, G.GIMRPHENOMD)
    op, oc = O.fourier_waveform(f, o, O.GIMRPHENOMD)
    wf_close(hp, op)
    wf_close(hc, oc)
    gp, _ = G.fourier_waveform(f, g, G.IMRPHENOMD)
    assert not np.array_equal(gp, hp)
    # IMRPhenomD ignores the gIMR fields
    gr, _ = G.fourier_waveform(f, G.make_params(**case, **EXTR), G.IMRPHENOMD)
    assert np.array_equal(gr, gp)


SINGLE_GIMR = ([("gimr_dphi", i + 4, 14, 2e-4 if i < 0 else 0.05) for i in range(-4, 10)] +
               [("gimr_dsigma", i, 3, 0.2) for i in range(3)] + [("gimr_dbeta", i, 2, 0.1) for i in range(2)] +
               [("gimr_dalpha", i, 4, 0.1) for i in range(4)])


@pytest.mark.parametrize("model", [G.GIMRPHENOMD, G.GIMRPHENOMPV2])
@pytest.mark.parametrize("field,idx,n,delta", SINGLE_GIMR)
def test_gimr_every_single_coefficient_matches_oracle(model, field, idx, n, delta):
    """each phi order -4..9 (8 and 9: the log parts, connected with the GR log derivative as the reference
    does) and each sigma / beta / alpha index on its own"""
    f = W.uniform_grid(20., 1024., 0.25)
    d = [0.] * n
    d[idx] = delta
    case = dict(CASES[0]) if model == G.GIMRPHENOMD else dict(mass1=36, mass2=29, spin1=[.3, .1, .2],
                                                               spin2=[-.1, .2, .1])
    g, o = both(**case, **EXTR, **{field: d})
    hp, hc = G.fourier_waveform(f, g, model)
    op, oc = O.fourier_waveform(f, o, model)
    wf_close(hp, op)
    wf_close(hc, oc)


def test_gimr_network_loglike_batch():
    f, psd, dets, data = _net()
    ev = G.Event(f, data, [psd] * 3, dets)
    P_ = W.perturbed_params(W.injection_params(1)[0], 24, step=5)
    rng = np.random.default_rng(12)
    P_["gimr_dphi"][:, 6] = rng.normal(0, 0.02, len(P_))
    P_["gimr_dphi"][:, 2] = rng.normal(0, 1e-4, len(P_))
    P_["gimr_dbeta"][:, 0] = rng.normal(0, 0.05, len(P_))
    P_["gimr_dalpha"][:, 3] = rng.normal(0, 0.05, len(P_))
    ll = G.loglike_batch(ev, P_, G.GIMRPHENOMD, G.FP64)
    for i in range(0, len(P_), 4):
        rec = P_[i]
        kw = {k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names}
        ref = O.network_loglike(f, data, [psd] * 3, dets, O.make_params(**kw), O.GIMRPHENOMD)
        assert abs(ll[i] - ref) <= LL_TOL * max(abs(ref), 1), (i, ll[i], ref)
    mx = G.loglike_batch(ev, P_, G.GIMRPHENOMD, G.MIXED)
    assert np.max(np.abs(mx - ll)) <= 1e-3
    ev.close()


def test_gimr_through_gwatpy_lists():
    f = W.uniform_grid(20., 512., 0.25)
    p = P.gen_params(mass1=36, mass2=29, spin1=(0, 0, .3), spin2=(0, 0, -.1), Luminosity_Distance=410.,
                     incl_angle=2.9, RA=1.37, DEC=-1.26, psi=2.66, gmst=W.GMST, tc=0.01, phiRef=0.3, f_ref=20.,
                     phii=[2, -2], delta_phi=[0.05, -1e-4], betai=[3], delta_beta=[0.1])
    st, hp, hc = P.fourier_waveform(f, "gIMRPhenomD", p)
    assert st == G.GWH_OK
    dphi = [0.] * 14
    dphi[2 + 4], dphi[-2 + 4] = 0.05, -1e-4
    o = O.make_params(**CASES[0], **EXTR, gimr_dphi=dphi, gimr_dbeta=[0., 0.1])
    op, oc = O.fourier_waveform(f, o, O.GIMRPHENOMD)
    wf_close(hp, op)
    # a gIMR parameter vector (fisher.cpp:2473-2505) written into the handle: the deltas follow the GR basis in
    # list order (phi 2, phi -2, beta 3)
    v, _ = O.unpack_parameters(o, "IMRPhenomD", 11)
    v = np.concatenate([v, [0.07, -2e-4, 0.2]])
    P.repack_parameters(v, p, "gIMRPhenomD", 14)
    o2 = O.make_params(**{**CASES[0], **EXTR, "gimr_nlist": 3, "gimr_list": [6, 2, 18]})
    o2 = O.repack_parameters(v, "gIMRPhenomD", 14, o2)
    assert o2.gimr_dphi[6] == 0.07 and o2.gimr_dphi[2] == -2e-4 and o2.gimr_dbeta[1] == 0.2
    st, hp, hc = P.fourier_waveform(f, "gIMRPhenomD", p)
    assert st == G.GWH_OK
    wf_close(hp, O.fourier_waveform(f, o2, O.GIMRPHENOMD)[0])
    p.close()


# ---- the modifications on IMRPhenomPv2 (ppE_IMRPhenomP.cpp, gIMRPhenomP.cpp) ---------------------------------
PV2_CASES = [dict(mass1=36, mass2=29, spin1=[0.3, 0.1, 0.2], spin2=[-0.1, 0.2, 0.1]),
             dict(mass1=50, mass2=12, spin1=[0.5, -0.3, -0.4], spin2=[0.1, 0.0, 0.3])]
PV2_MODS = [(G.PPE_IMRPHENOMPV2_INSPIRAL, TERMS[1]), (G.PPE_IMRPHENOMPV2_IMR, TERMS[2]),
            (G.GIMRPHENOMPV2, GIMR[0]), (G.GIMRPHENOMPV2, GIMR[1])]


@pytest.mark.parametrize("case", PV2_CASES)
@pytest.mark.parametrize("model,mods", PV2_MODS)
def test_pv2_variants_waveform_matches_oracle(case, model, mods):
    f = W.uniform_grid(20., 1024., 0.25)
    g, o = both(**case, **EXTR, **mods)
    hp, hc = G.fourier_waveform(f, g, model)
    op, oc = O.fourier_waveform(f, o, model)
    wf_close(hp, op)
    wf_close(hc, oc)
    gp, _ = G.fourier_waveform(f, g, G.IMRPHENOMPV2)
    assert not np.array_equal(gp, hp)
    # with zero modifications the variant is IMRPhenomPv2 bit for bit
    z, _ = G.fourier_waveform(f, G.make_params(**case, **EXTR), model)
    gz, _ = G.fourier_waveform(f, G.make_params(**case, **EXTR), G.IMRPHENOMPV2)
    assert np.array_equal(z, gz)


@pytest.mark.parametrize("model", [G.PPE_IMRPHENOMPV2_IMR, G.GIMRPHENOMPV2])
def test_pv2_variants_network_loglike_batch(model):
    f, psd, dets, data = _net()
    ev = G.Event(f, data, [psd] * 3, dets)
    P_ = W.precessing_params(24)
    rng = np.random.default_rng(13)
    if model == G.PPE_IMRPHENOMPV2_IMR:
        P_["ppe_nmod"] = 1
        P_["ppe_b"][:, 0] = -7.
        P_["ppe_beta"][:, 0] = rng.normal(0, 3e-5, len(P_))
    else:
        P_["gimr_dphi"][:, 6] = rng.normal(0, 0.02, len(P_))
        P_["gimr_dalpha"][:, 3] = rng.normal(0, 0.05, len(P_))
    ll = G.loglike_batch(ev, P_, model, G.FP64)
    for i in range(0, len(P_), 4):
        rec = P_[i]
        kw = {k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names}
        ref = O.network_loglike(f, data, [psd] * 3, dets, O.make_params(**kw), model)
        assert abs(ll[i] - ref) <= LL_TOL * max(abs(ref), 1), (i, ll[i], ref)
    mx = G.loglike_batch(ev, P_, model, G.MIXED)
    assert np.max(np.abs(mx - ll)) <= 1e-3
    for i in (0, 7):
        assert G.loglike_batch(ev, P_[i:i + 1], model, G.FP64)[0] == ll[i]
    ev.close()


def test_pv2_variants_through_gwatpy():
    f = W.uniform_grid(20., 512., 0.25)
    kw = dict(mass1=36, mass2=29, Luminosity_Distance=410., incl_angle=2.9, RA=1.37, DEC=-1.26, psi=2.66,
              gmst=W.GMST, tc=0.01, phiRef=0.3, f_ref=20.)
    p = P.gen_params(spin1=(.3, .1, .2), spin2=(-.1, .2, .1), Nmod=2, bppe=[-7., -3.], betappe=[3e-5, -3e-3],
                     phii=[2], delta_phi=[0.05], **kw)
    st, hp, hc = P.fourier_waveform(f, "ppE_IMRPhenomPv2_Inspiral", p)
    assert st == G.GWH_OK
    base = dict(PV2_CASES[0], **EXTR)
    op, oc = O.fourier_waveform(f, O.make_params(**base, **TERMS[1]), O.PPE_IMRPHENOMPV2_INSPIRAL)
    wf_close(hp, op)
    wf_close(hc, oc)
    st, hp, hc = P.fourier_waveform(f, "gIMRPhenomPv2", p)
    assert st == G.GWH_OK
    dphi = [0.] * 14
    dphi[2 + 4] = 0.05
    op, oc = O.fourier_waveform(f, O.make_params(**base, gimr_dphi=dphi), O.GIMRPHENOMPV2)
    wf_close(hp, op)
    p.close()


# ---- Fisher matrices with the ppE betas in the basis (fisher.cpp:42-60, :2083-2088, :2424-2429) --------------
FISHER_TOL = 1e-6
EPS = 1e-5  # the well-conditioned step of the GPU Fisher parity tests


def _fnorm(a, b):
    s = np.sqrt(np.abs(np.diag(b)))
    return np.max(np.abs(a - b) / np.outer(s, s)), np.max(np.abs(np.diag(a) - np.diag(b)) / np.abs(np.diag(b)))


FISHER_CASES = [  # (model, engine basis, oracle method, GR dim, sky, ppE terms)
    (G.PPE_IMRPHENOMD_IMR, G.BASIS_PHENOMD_FULL, "ppE_IMRPhenomD_IMR", 11, 0, TERMS[1]),
    (G.PPE_IMRPHENOMD_INSPIRAL, G.BASIS_PHENOMD_FULL, "ppE_IMRPhenomD_Inspiral", 11, 0, TERMS[0]),
    (G.PPE_IMRPHENOMD_INSPIRAL, G.BASIS_PHENOMD_SKYAVG, "ppE_IMRPhenomD_Inspiral", 7, 1, TERMS[1]),
    (G.PPE_IMRPHENOMD_IMR, G.BASIS_PHENOMD_MCMC, "MCMC_ppE_IMRPhenomD_IMR", 11, 0, TERMS[2]),
    (G.PPE_IMRPHENOMPV2_IMR, G.BASIS_PV2_FULL, "ppE_IMRPhenomPv2_IMR", 13, 0, TERMS[1]),
]


@pytest.mark.parametrize("model,basis,method,gdim,sky,terms", FISHER_CASES)
@pytest.mark.parametrize("order", [2, 4])
def test_ppe_fisher_vs_oracle(model, basis, method, gdim, sky, terms, order):
    f = W.uniform_grid(20., 1024., 0.25)
    psd = W.analytic_psd(f)
    kw = dict(mass1=36., mass2=29., spin1=[0, 0, 0.3], spin2=[0, 0, -0.1], DL=410., tc=0., phiRef=0., f_ref=20.,
              RA=1.37, DEC=-1.26, psi=2.66, incl=2.9, gmst=W.GMST, sky_average=sky, **terms)
    if gwh_pv2(model):
        kw.update(spin1=[0.3, 0.1, 0.3], spin2=[-0.1, 0.2, -0.1], chip=0.35, phip=0.8)
    dim = gdim + terms["ppe_nmod"]
    gp, op = G.make_params(**kw), O.make_params(**kw)
    a = G.fisher_numerical(f, psd, basis, 0, 0, dim, gp, order=order, epsilon=EPS, model=model)
    b = O.fisher_numerical(f, psd, method, 0, 0, dim, op, order=order, epsilon=EPS)
    e, ed = _fnorm(a, b)
    assert e < FISHER_TOL and ed < FISHER_TOL, (e, ed)
    assert np.allclose(a, a.T, rtol=0, atol=0)


# gIMR bases: phi orders -2 and 2, sigma_3, beta_2, alpha_4 (GWH_GIMR_CODE_*), at nonzero deltas; a repeat and an
# unread index (-1) give zero rows
GIMR_FISHER = dict(gimr_nlist=5, gimr_list=[2, 6, 15, 17, 21], gimr_dphi=[0, 0, 2e-4] + [0] * 3 + [0.05] + [0] * 7,
                   gimr_dsigma=[0, 0.1, 0], gimr_dbeta=[-0.05, 0], gimr_dalpha=[0, 0, 0.1, 0])
GIMR_FISHER_CASES = [  # (model, engine basis, oracle method, GR dim, sky, list (None: the 5 above), zero rows)
    (G.GIMRPHENOMD, G.BASIS_PHENOMD_FULL, "gIMRPhenomD", 11, 0, None, []),
    (G.GIMRPHENOMD, G.BASIS_PHENOMD_SKYAVG, "gIMRPhenomD", 7, 1, None, []),
    (G.GIMRPHENOMD, G.BASIS_PHENOMD_MCMC, "MCMC_gIMRPhenomD", 11, 0, [6, -1, 6, 21], [12, 13]),
    (G.GIMRPHENOMPV2, G.BASIS_PV2_FULL, "gIMRPhenomPv2", 13, 0, [6, 15, 21], []),  # 16 parameters at most
]


@pytest.mark.parametrize("model,basis,method,gdim,sky,lst,zero", GIMR_FISHER_CASES)
def test_gimr_fisher_vs_oracle(model, basis, method, gdim, sky, lst, zero):
    f = W.uniform_grid(20., 1024., 0.25)
    psd = W.analytic_psd(f)
    kw = dict(mass1=36., mass2=29., spin1=[0, 0, 0.3], spin2=[0, 0, -0.1], DL=410., tc=0., phiRef=0., f_ref=20.,
              RA=1.37, DEC=-1.26, psi=2.66, incl=2.9, gmst=W.GMST, sky_average=sky, **GIMR_FISHER)
    if lst is not None:
        kw.update(gimr_nlist=len(lst), gimr_list=lst)
    if gwh_pv2(model):
        kw.update(spin1=[0.3, 0.1, 0.3], spin2=[-0.1, 0.2, -0.1], chip=0.35, phip=0.8)
    dim = gdim + kw["gimr_nlist"]
    gp, op = G.make_params(**kw), O.make_params(**kw)
    a = G.fisher_numerical(f, psd, basis, 0, 0, dim, gp, order=2, epsilon=EPS, model=model)
    b = O.fisher_numerical(f, psd, method, 0, 0, dim, op, order=2, epsilon=EPS)
    # the zero rows of a repeat / an unread index: the two stencil waveforms are equal, and the fused
    # multiply-adds of the stencil sum (c h - c h) leave only the rounding of c h
    keep = np.diag(b) != 0
    assert np.all(b[~keep] == 0) and np.max(np.abs(a[~keep]), initial=0) <= 1e-12 * np.max(np.diag(b))
    e, ed = _fnorm(a[np.ix_(keep, keep)], b[np.ix_(keep, keep)])
    assert e < FISHER_TOL and ed < FISHER_TOL, (e, ed)
    assert np.allclose(a, a.T, rtol=0, atol=0)
    assert list(np.flatnonzero(~keep)) == zero


def gwh_pv2(model):
    return model in (G.IMRPHENOMPV2, G.PPE_IMRPHENOMPV2_INSPIRAL, G.PPE_IMRPHENOMPV2_IMR, G.GIMRPHENOMPV2)


def test_ppe_fisher_batch_and_two_detectors():
    """the batched Fisher of ppE points on a resident network event against one-point Fishers summed over the
    detectors; and one point on (H1 relative to L1) against the oracle"""
    f = W.uniform_grid(20., 512., 0.25)
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    ev = G.Event(f, [np.zeros(len(f), complex)] * 3, [psd] * 3, dets)
    P_ = W.perturbed_params(W.injection_params(1)[0], 6, step=3)
    P_["ppe_nmod"] = 2
    P_["ppe_b"][:, 0], P_["ppe_b"][:, 1] = -7., -3.
    P_["ppe_beta"][:, 0] = np.linspace(-2e-5, 2e-5, len(P_))
    F, sig = G.fisher_batch(ev, P_, G.PPE_IMRPHENOMD_IMR, G.BASIS_PHENOMD_FULL, 13, epsilon=EPS, with_sigma=True)
    assert F.shape == (len(P_), 13, 13) and np.all(np.isfinite(sig))
    for i in (0, 4):
        rec = P_[i]
        kw = {k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names}
        op = O.make_params(**kw)
        ref = sum(O.fisher_numerical(f, psd, "ppE_IMRPhenomD_IMR", d, dets[0], 13, op, epsilon=EPS) for d in dets)
        e, ed = _fnorm(F[i], ref)
        assert e < FISHER_TOL and ed < FISHER_TOL, (i, e, ed)
    ev.close()


# ---- theories mapped onto the ppE terms (ppE_utilities.cpp:158-360) ------------------------------------------
THEORIES = [  # (theory, theory parameters, exponents (PNSeries / ppEAlt), IMRPhenomD model)
    (G.THEORY_DCS, [1e-17], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_EDGB, [1e-20], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_EDGB_HO_LO, [1e-20], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_EDGB_GHOV1, [1e-20, 1.5], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_EDGB_GHOV2, [1e-20, 1.5], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_EDGB_GHOV3, [1e-20, 1.5], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_EXTRADIMENSION, [1e-12], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_BHEVAPORATION, [1e-6], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_TVG, [1e-3], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_DIPRAD, [1e-4], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_NONCOMM, [0.05], [], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_PNSERIES, [1e-3, 0.5], [-7., -5.], G.PPE_IMRPHENOMD_IMR),
    (G.THEORY_PPEALT, [3e-5, -3e-3], [-7., -3.], G.PPE_IMRPHENOMD_INSPIRAL),
    (G.THEORY_MODDISPERSION, [3e-15], [4.5], G.PPE_IMRPHENOMD_IMR),
]
TO_PV2 = {G.PPE_IMRPHENOMD_INSPIRAL: G.PPE_IMRPHENOMPV2_INSPIRAL, G.PPE_IMRPHENOMD_IMR: G.PPE_IMRPHENOMPV2_IMR}


def _theory_kw(th, beta, b, **extra):
    return dict(ppe_theory=th, ppe_nmod=len(beta), ppe_beta=beta, ppe_b=b or [0.] * len(beta), **extra)


@pytest.mark.parametrize("th,beta,b,model", THEORIES)
@pytest.mark.parametrize("pv2", [False, True])
def test_theory_waveform_matches_oracle(th, beta, b, model, pv2):
    f = W.uniform_grid(20., 1024., 0.25)
    kw = dict(CASES[0], **EXTR, **_theory_kw(th, beta, b))
    if pv2:
        model = TO_PV2[model]
        kw.update(spin1=[0.3, 0.1, 0.3], spin2=[-0.1, 0.2, -0.1], chip=0.35, phip=0.8)
    g, o = both(**kw)
    hp, hc = G.fourier_waveform(f, g, model)
    op, oc = O.fourier_waveform(f, o, model)
    wf_close(hp, op)
    wf_close(hc, oc)
    gr, _ = G.fourier_waveform(f, g, G.IMRPHENOMPV2 if pv2 else G.IMRPHENOMD)
    nz = np.abs(op) > 0
    assert np.max(np.abs(np.angle(hp[nz] / gr[nz]))) > 1e-3  # the theory term is visible


def test_theory_network_loglike_batch_and_refusal():
    """one batch of instances with different theories (and cosmologies), FP64 against the oracle, mixed within
    1e-3; an instance whose theory cannot be mapped (a GHO variant given one parameter) is NaN, the others not"""
    f, psd, dets, data = _net()
    ev = G.Event(f, data, [psd] * 3, dets)
    ins = [t for t in THEORIES if t[3] == G.PPE_IMRPHENOMD_INSPIRAL]
    P_ = W.perturbed_params(W.injection_params(1)[0], 4 * len(ins), step=5)
    for i in range(len(P_)):
        th, beta, b, _ = ins[i % len(ins)]
        P_["ppe_theory"][i], P_["ppe_nmod"][i] = th, len(beta)
        P_["ppe_beta"][i, :len(beta)] = beta
        P_["ppe_b"][i, :len(b)] = b
        P_["cosmology"][i] = (i // len(ins)) % 5
    bad = 3  # EDGB_GHOV1 with a single parameter
    assert P_["ppe_theory"][bad] == G.THEORY_EDGB_GHOV1
    P_["ppe_nmod"][bad] = 1
    ll = G.loglike_batch(ev, P_, G.PPE_IMRPHENOMD_INSPIRAL, G.FP64)
    assert np.isnan(ll[bad]) and np.all(np.isfinite(np.delete(ll, bad)))
    for i in range(len(P_)):
        if i == bad:
            continue
        rec = P_[i]
        kw = {k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names}
        ref = O.network_loglike(f, data, [psd] * 3, dets, O.make_params(**kw), G.PPE_IMRPHENOMD_INSPIRAL)
        assert abs(ll[i] - ref) <= LL_TOL * max(abs(ref), 1), (i, ll[i], ref)
    mx = G.loglike_batch(ev, P_, G.PPE_IMRPHENOMD_INSPIRAL, G.MIXED)
    ok = np.isfinite(ll)
    assert np.isnan(mx[bad]) and np.max(np.abs(mx[ok] - ll[ok])) <= 1e-3
    ev.close()


@pytest.mark.parametrize("model,basis,gdim,th,val", [
    (G.PPE_IMRPHENOMD_INSPIRAL, G.BASIS_PHENOMD_FULL, 11, G.THEORY_NONCOMM, 0.05),
    (G.PPE_IMRPHENOMPV2_INSPIRAL, G.BASIS_PV2_FULL, 13, G.THEORY_DIPRAD, 1e-4)])
def test_theory_fisher_vs_oracle(model, basis, gdim, th, val):
    """a theory's parameter appended to the GR basis as betappe (fisher.cpp:52-64), the theory map in every
    perturbed record"""
    f = W.uniform_grid(20., 1024., 0.25)
    psd = W.analytic_psd(f)
    kw = dict(mass1=36., mass2=29., spin1=[0, 0, 0.3], spin2=[0, 0, -0.1], DL=410., tc=0., phiRef=0., f_ref=20.,
              RA=1.37, DEC=-1.26, psi=2.66, incl=2.9, gmst=W.GMST, ppe_theory=th, ppe_nmod=1, ppe_beta=[val],
              cosmology=G.COSMO_WMAP9)
    method = "ppE_IMRPhenomD_Inspiral"
    if model == G.PPE_IMRPHENOMPV2_INSPIRAL:
        kw.update(spin1=[0.3, 0.1, 0.3], spin2=[-0.1, 0.2, -0.1], chip=0.35, phip=0.8)
        method = "ppE_IMRPhenomPv2_Inspiral"
    gp, op = G.make_params(**kw), O.make_params(**kw)
    a = G.fisher_numerical(f, psd, basis, 0, 0, gdim + 1, gp, order=2, epsilon=EPS, model=model)
    b = O.fisher_numerical(f, psd, method, 0, 0, gdim + 1, op, order=2, epsilon=EPS)
    e, ed = _fnorm(a, b)
    assert e < FISHER_TOL and ed < FISHER_TOL, (e, ed)


def test_theory_method_strings_through_gwatpy():
    """"dCS_IMRPhenomD" and the like name the theory and, through the mapping, the ppE method; the handle's
    cosmology string picks the redshift fit"""
    f = W.uniform_grid(20., 512., 0.25)
    for method, th, beta, cosmo, model in [("dCS_IMRPhenomD", G.THEORY_DCS, [1e-17], "PLANCK15",
                                            G.PPE_IMRPHENOMD_INSPIRAL),
                                           ("EdGB_HO_IMRPhenomD", G.THEORY_EDGB, [1e-20], "wmap9",
                                            G.PPE_IMRPHENOMD_INSPIRAL),
                                           ("EdGB_GHOv3_IMRPhenomPv2", G.THEORY_EDGB_GHOV3, [1e-20, 1.5], "WMAP7",
                                            G.PPE_IMRPHENOMPV2_INSPIRAL),
                                           ("TVG_IMRPhenomD", G.THEORY_TVG, [1e-3], "Planck13",
                                            G.PPE_IMRPHENOMD_INSPIRAL)]:
        p = P.gen_params(mass1=36, mass2=29, spin1=(0, 0, .3), spin2=(0, 0, -.1), Luminosity_Distance=410.,
                         incl_angle=2.9, RA=1.37, DEC=-1.26, psi=2.66, gmst=W.GMST, tc=0.01, phiRef=0.3, f_ref=20.,
                         cosmology=cosmo, Nmod=len(beta), betappe=beta)
        st, hp, hc = P.fourier_waveform(f, method, p)
        assert st == G.GWH_OK, method
        o = O.make_params(**CASES[0], **EXTR, ppe_theory=th, ppe_nmod=len(beta), ppe_beta=beta,
                          cosmology=["PLANCK15", "PLANCK13", "WMAP9", "WMAP7", "WMAP5"].index(cosmo.upper()))
        wf_close(hp, O.fourier_waveform(f, o, model)[0])
        p.close()
    p = P.gen_params(mass1=36, mass2=29, Nmod=1, betappe=[1.])
    st, _, _ = P.fourier_waveform(f, "polarization_test_IMRPhenomD", p)
    assert st == G.GWH_ERR_UNSUPPORTED
    p.close()
