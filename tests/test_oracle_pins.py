"""Pins of the CPU oracle against the reference's own fixtures and against what
the mathematics fixes.  A plausible transcription mistake (dropped term, wrong
sign, wrong index) in the oracle fails at least one of these."""
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fisher_data_subset.npz")


def fixture_grid():
    # testing/fisher_comparison.cpp:51-62
    length, flow, fhigh = 3000, 15., 1000.
    df = (fhigh - flow) / (length - 1)
    f = np.array([flow + i * df for i in range(length)])
    return f, O.hanford_o1_fitted_asd(f) ** 2


def fixture_params(row, **kw):
    m1, m2, s1, s2, DL, tc, phic = row
    base = dict(mass1=m1, mass2=m2, spin1=[0, 0, s1], spin2=[0, 0, s2], DL=DL, tc=tc, phiRef=phic,
                sky_average=1, shift_time=0, shift_phase=0, dep_postmerger=1)
    base.update(kw)
    return O.make_params(**base)


def schur(F, E=(0, 1, 2), I=(3, 4, 5, 6)):
    E, I = list(E), list(I)
    return F[np.ix_(I, I)] - F[np.ix_(E, I)].T @ np.linalg.solve(F[np.ix_(E, E)], F[np.ix_(E, I)])


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_fisher_fixture_amplitude_elements(gold):
    """ln A0, phi_c, t_c elements depend only on the amplitude model and PSD.
    The fixtures predate the QNM spline (they use the closed-form ringdown fits,
    dep_postmerger) and the LAL-digit calibration table; with the current table the
    residual is <= 5e-4 (<= 6e-5 with the paper's rounded table)."""
    f, psd = fixture_grid()
    par, aut = gold["parameters"], gold["autodiff"]
    worst = 0
    for r in range(0, 40, 2):
        F = O.fisher_numerical(f, psd, "IMRPhenomD", 0, 0, 7, fixture_params(par[r]))
        Fa = aut[r]
        for (i, j) in [(0, 0), (1, 1), (2, 2)]:
            worst = max(worst, abs(F[i, j] / Fa[i, j] - 1))
        # (lnA0, tc) and (lnA0, phi) vanish identically
        assert F[0, 1] == 0 and F[0, 2] == 0
        # phi_c - t_c cross term: sign flipped by the reference's 2023-05-21 phase convention change
        assert abs(F[1, 2] / (-Fa[1, 2]) - 1) < 2e-3
    assert worst < 5e-4, worst


def test_fisher_fixture_profiled_intrinsic_block(gold):
    """The intrinsic block profiled over (ln A0, phi_c, t_c) is invariant under the
    reference's tc/phic convention changes; it pins the phase model and its
    chirp-mass, eta and spin derivatives to the level of the reference's own
    numerical-vs-autodiff spread."""
    f, psd = fixture_grid()
    par, aut, num = gold["parameters"], gold["autodiff"], gold["numerical"]
    for r in range(0, 30, 3):
        F = O.fisher_numerical(f, psd, "IMRPhenomD", 0, 0, 7, fixture_params(par[r]))
        Sm, Sa = schur(F), schur(aut[r])
        scale = np.sqrt(np.outer(np.diag(Sa), np.diag(Sa)))
        err = np.max(np.abs(Sm - Sa) / scale)
        ref_spread = np.max(np.abs(schur(num[r]) - Sa) / scale)
        assert err < max(3e-3, 5 * ref_spread), (r, err, ref_spread)


def test_fisher_fixture_catches_planted_error(gold):
    """The same comparison fails when a term is dropped (amplitude scaled by the
    inclination factor that only the non-averaged path applies)."""
    f, psd = fixture_grid()
    par, aut = gold["parameters"], gold["autodiff"]
    p = fixture_params(par[0])
    amp = O.fourier_amplitude(f, p) * (1 + np.cos(0.3) ** 2) / 2
    hh = 4 * O.simpsons_sum(f[1] - f[0], amp ** 2 / psd)
    assert abs(hh / aut[0][0, 0] - 1) > 1e-2


def test_spline_matches_textbook_natural_spline():
    rng = np.random.default_rng(1)
    x = np.sort(rng.uniform(-1, 1, 40))
    y = np.sin(3 * x) + 0.1 * rng.normal(size=40)
    cs = CubicSpline(x, y, bc_type="natural")
    for xq in np.linspace(x[0], x[-1], 57):
        assert abs(O.spline_eval(x, y, xq, 0) - cs(xq)) < 1e-12
        assert abs(O.spline_eval(x, y, xq, 1) - cs(xq, 1)) < 1e-10
    assert np.isnan(O.spline_eval(x, y, x[-1] + 1e-3, 0))


@pytest.mark.parametrize("m1,m2,s1,s2", [(36, 29, 0.3, -0.1), (80, 10, -0.8, 0.5), (5, 4, 0.9, 0.9),
                                         (150, 140, -0.95, -0.95)])
def test_phenomd_c1_connections(m1, m2, s1, s2):
    """Amplitude C1 at f1, f3 and phase C1 at f1_phase, f2_phase (by construction of
    the connection coefficients)."""
    p = O.make_params(mass1=m1, mass2=m2, spin1=[0, 0, s1], spin2=[0, 0, s2], DL=400)
    s = O.phenomd_setup_summary(p)
    a = O.phenomd_pieces(p, s["f1"])
    assert abs(a["amp_ins"] / a["amp_int"] - 1) < 1e-10
    a = O.phenomd_pieces(p, s["f3"])
    assert abs(a["amp_mr"] / a["amp_int"] - 1) < 1e-9
    h = 1e-7 * s["f3"]
    up, dn = O.phenomd_pieces(p, s["f3"] + h), O.phenomd_pieces(p, s["f3"] - h)
    d_int = (up["amp_int"] - dn["amp_int"]) / (2 * h)
    assert abs(d_int - a["damp_mr"]) < 1e-5 * abs(a["amp_mr"]) / s["f3"]
    b = O.phenomd_pieces(p, s["f1_phase"])
    assert abs(b["phase_ins"] - b["phase_int"]) < 1e-9 * abs(b["phase_ins"])
    assert abs(b["dphase_ins"] - b["dphase_int"]) < 1e-9 * abs(b["dphase_ins"])
    c = O.phenomd_pieces(p, s["f2_phase"])
    assert abs(c["phase_int"] - c["phase_mr"]) < 1e-9 * max(1, abs(c["phase_int"]))
    assert abs(c["dphase_int"] - c["dphase_mr"]) < 1e-9 * abs(c["dphase_int"])
    # analytic frequency derivatives agree with central differences
    f0 = 0.7 * s["f1"]
    h = 1e-6 * f0
    up, dn, mid = O.phenomd_pieces(p, f0 + h), O.phenomd_pieces(p, f0 - h), O.phenomd_pieces(p, f0)
    assert abs((up["phase_ins"] - dn["phase_ins"]) / (2 * h) - mid["dphase_ins"]) < 1e-6 * abs(mid["dphase_ins"])
    assert abs((up["amp_ins"] - dn["amp_ins"]) / (2 * h) - mid["damp_ins"]) < 1e-6 * abs(mid["damp_ins"])


def test_zero_above_cutoff_and_polarizations():
    p = O.make_params(mass1=36, mass2=29, spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410, incl=2.9, f_ref=20)
    f = np.linspace(20, 1024, 4017)
    hp, hc = O.fourier_waveform(f, p)
    fc = O.fcut(p)
    assert np.all(hp[f > fc] == 0) and np.all(hc[f > fc] == 0)
    assert np.all(np.abs(hp[f <= fc]) > 0)
    ci = np.cos(2.9)
    # h_x = -i h_+ cos(i) / ((1+cos^2 i)/2)
    np.testing.assert_allclose(hc, -1j * hp * ci / ((1 + ci * ci) / 2), rtol=1e-13, atol=0)


def test_simpson_rule():
    # exact for cubics with an odd number of points
    x = np.linspace(0, 2, 101)
    y = 3 * x ** 3 - x ** 2 + 2
    assert abs(O.simpsons_sum(x[1] - x[0], y) - (3 * 16 / 4 - 8 / 3 + 4)) < 1e-12
    # weights follow index parity whatever n is (even n is not Simpson's rule)
    for n in range(3, 10):
        w = O.simpsons_sum(1.0, np.ones(n)) * 3
        expect = 2 + sum(4 if i % 2 else 2 for i in range(1, n - 1))
        assert abs(w - expect) < 1e-12


def test_antenna_invariants():
    rng = np.random.default_rng(7)
    for _ in range(50):
        ra, dec, psi, gmst = rng.uniform(0, 2 * np.pi), rng.uniform(-1.5, 1.5), rng.uniform(0, np.pi), rng.uniform(0, 6)
        for det in range(3):
            fp, fx = O.antenna(det, ra, dec, psi, gmst)
            fp2, fx2 = O.antenna(det, ra, dec, psi + 0.37, gmst)
            assert abs(fp * fp + fx * fx - fp2 * fp2 - fx2 * fx2) < 1e-12
            assert fp * fp + fx * fx <= 1 + 1e-12
            fp3, _ = O.antenna(det, ra, dec, psi + np.pi / 4, gmst)
            assert abs(fp3 - fx) < 1e-12
        d01 = O.dtoa(ra, dec, gmst, 0, 1)
        assert abs(d01 + O.dtoa(ra, dec, gmst, 1, 0)) < 1e-18
        assert abs(d01) <= 3002e3 / 299792458.


def test_loglike_identity_data_equals_signal():
    """logL(d = h) = (h|h)/2 = SNR^2/2 (tests/src/test_mcmc.cpp:1064-1071)."""
    f = np.arange(20, 1024.01, 0.25)
    x = f / 215.
    psd = 1e-49 * (x ** -4.14 - 5 * x ** -2 + 111 * (1 - x ** 2 + x ** 4 / 2) / (1 + x ** 2 / 2))
    p = O.make_params(mass1=36, mass2=29, spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410, incl=2.9, RA=1.37,
                      DEC=-1.26, psi=2.66, f_ref=20)
    r = O.detector_response(f, p, 0)
    ll = O.log_likelihood(r, psd, f, p, 0)
    snr = O.snr(psd, r, f)
    assert abs(ll - snr ** 2 / 2) < 1e-10 * ll
    # network version: each detector's data is its own time-shifted response
    dets = [0, 1, 2]
    data = []
    for d in dets:
        rd = O.detector_response(f, p, d)
        dt = O.dtoa(p.RA, p.DEC, p.gmst, dets[0], d)
        data.append(rd * np.exp(1j * (-2 * np.pi * dt) * f))
    psds = [psd] * 3
    ll3 = O.network_loglike(f, data, psds, dets, p)
    hh = sum(4 * O.simpsons_sum(f[len(f) // 2] - f[len(f) // 2 - 1], np.abs(dd) ** 2 / psd) for dd in data)
    assert abs(ll3 - hh / 2) < 1e-10 * abs(ll3)


def test_full_basis_fisher_closed_forms():
    """Full (non sky-averaged) IMRPhenomD basis: h is proportional to 1/DL, so
    F(lnDL, lnDL) = (h|h); phi_ref enters as 2 phi_ref, so F(phi,phi) = 4 (h|h)."""
    f = np.arange(20, 1024.01, 0.25)
    x = f / 215.
    psd = 1e-49 * (x ** -4.14 - 5 * x ** -2 + 111 * (1 - x ** 2 + x ** 4 / 2) / (1 + x ** 2 / 2))
    p = O.make_params(mass1=36, mass2=29, spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410, incl=2.9, RA=1.37,
                      DEC=-1.26, psi=2.66, f_ref=20)
    F = O.fisher_numerical(f, psd, "IMRPhenomD", 0, 0, 11, p)
    r = O.detector_response(f, p, 0)
    hh = 4 * O.simpsons_sum(f[1] - f[0], np.abs(r) ** 2 / psd)
    assert abs(F[6, 6] / hh - 1) < 1e-5
    assert abs(F[4, 4] / (4 * hh) - 1) < 1e-5
    # t_c enters as 2 pi t_c (f - f_ref) with shift_time forced off:
    # F(tc,tc) = 4 (2 pi)^2 ((f - f_ref) h | (f - f_ref) h)
    ftt = 4 * O.simpsons_sum(f[1] - f[0], (2 * np.pi * (f - 20.)) ** 2 * np.abs(r) ** 2 / psd)
    assert abs(F[5, 5] / ftt - 1) < 1e-4
    # one detector leaves the sky directions nearly degenerate; the inverse is
    # checked on a well-conditioned symmetric matrix of the same size instead
    rng = np.random.default_rng(3)
    B = rng.normal(size=(11, 11))
    S = B @ B.T + 11 * np.eye(11)
    np.testing.assert_allclose(S @ O.lu_invert(S), np.eye(11), atol=1e-12)


def test_pv2_aligned_limit_matches_phenomd_power():
    """With no in-plane spin the twist-up reduces to the (2,+-2) modes of IMRPhenomD:
    |h+|^2 + |hx|^2 is identical (phase conventions aside)."""
    f = np.arange(20, 1024.01, 0.25)
    kw = dict(mass1=36, mass2=29, spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410, incl=1.1, f_ref=20)
    hp, hc = O.fourier_waveform(f, O.make_params(**kw), O.IMRPHENOMD)
    qp, qc = O.fourier_waveform(f, O.make_params(**kw), O.IMRPHENOMPV2)
    a = np.abs(hp) ** 2 + np.abs(hc) ** 2
    b = np.abs(qp) ** 2 + np.abs(qc) ** 2
    np.testing.assert_allclose(b, a, rtol=1e-9, atol=1e-60)
    s = O.pv2_setup_summary(O.make_params(**kw))
    assert abs(s["thetaJN"] - 1.1) < 1e-12 and abs(s["chip"]) < 1e-15


def test_pv2_precessing_is_finite_and_continuous():
    f = np.arange(20, 1024.01, 0.25)
    p = O.make_params(mass1=36, mass2=29, spin1=[0.5, 0.3, .3], spin2=[-0.2, 0.4, -.1], DL=410, incl=1.1,
                      f_ref=20, RA=1.0, DEC=0.3, psi=0.2)
    hp, hc = O.fourier_waveform(f, p, O.IMRPHENOMPV2)
    assert np.all(np.isfinite(hp)) and np.all(np.isfinite(hc))
    s = O.pv2_setup_summary(p)
    assert 0 < s["chip"] < 1
    nz = f <= O.fcut(p)
    # bin to bin the modulus changes smoothly (no branch jumps in the Euler angles)
    a = np.abs(hp[nz]) ** 2 + np.abs(hc[nz]) ** 2
    rel = np.abs(np.diff(a)) / a[:-1]
    assert np.max(rel) < 0.1


def test_gmst_against_meeus_series():
    """gps_to_GMST_radian (util.cpp:1797-1846) against the IAU 1982 sidereal-time series (Meeus,
    Astronomical Algorithms, eq. 12.4) with UTC = GPS - leap seconds.  The reference treats GPS seconds
    as UT (no leap seconds), so agreement is to ~1e-3 rad; a wrong epoch, rate or hour wrap fails."""
    for gps, leap in [(1126259462.4, 17.), (1187008882.4, 18.), (630763213., 13.), (1300000000., 18.)]:
        jd = 2444244.5 + (gps - leap) / 86400.  # GPS epoch 1980-01-06 00:00 UTC
        T = (jd - 2451545.0) / 36525.
        deg = 280.46061837 + 360.98564736629 * (jd - 2451545.0) + 0.000387933 * T * T - T ** 3 / 38710000.
        ref = np.deg2rad(deg % 360.)
        got = O.gps_to_gmst_radian(gps)
        d = (got - ref + np.pi) % (2 * np.pi) - np.pi
        assert abs(d) < 2e-3, (gps, got, ref)


@pytest.mark.parametrize("method,dim,sky", [("IMRPhenomD", 7, 1), ("IMRPhenomD", 11, 0), ("MCMC_IMRPhenomD", 11, 0),
                                            ("IMRPhenomPv2", 13, 0), ("MCMC_IMRPhenomPv2", 15, 0)])
def test_unpack_repack_round_trip_and_closed_forms(method, dim, sky):
    """repack(unpack(p)) == p, and the basis entries equal their closed forms (chirp mass, eta, logs)"""
    kw = dict(mass1=36., mass2=29., spin1=[0.2, -0.1, 0.3], spin2=[0.05, 0.3, -0.1], DL=410., tc=0.01, phiRef=0.3,
              incl=2.9, RA=1.37, DEC=-1.26, psi=2.66, f_ref=20., sky_average=sky)
    if method == "IMRPhenomPv2":
        kw.update(chip=0.4, phip=1.2)
    p = O.make_params(**kw)
    v, lf = O.unpack_parameters(p, method, dim)
    mc = (36. * 29.) ** 0.6 / 65. ** 0.2
    eta = 36. * 29. / 65. ** 2
    if dim == 7:
        assert abs(v[3] - mc) < 1e-12 and abs(v[4] - eta) < 1e-15 and list(np.nonzero(lf)[0]) == [0, 3, 4]
    elif method.startswith("MCMC"):
        assert abs(v[6] - np.log(410.)) < 1e-14 and abs(v[7] - np.log(mc)) < 1e-14 and abs(v[1] - np.sin(-1.26)) < 1e-15
        assert not lf.any()
    else:
        assert abs(v[6] - 410.) < 1e-12 and abs(v[7] - mc) < 1e-12 and list(np.nonzero(lf)[0]) == [6, 7]
    q = O.repack_parameters(v, method, dim, p)
    for name in ("mass1", "mass2", "DL", "tc", "phiRef"):
        assert abs(getattr(q, name) - getattr(p, name)) <= 1e-10 * abs(getattr(p, name)) + 1e-14, name
    if dim > 7:
        for name in ("RA", "DEC", "psi", "incl"):
            assert abs(getattr(q, name) - getattr(p, name)) < 1e-12, name
    if method == "MCMC_IMRPhenomPv2":
        assert np.allclose(list(q.spin1), kw["spin1"], atol=1e-12) and np.allclose(list(q.spin2), kw["spin2"], atol=1e-12)


# ---- ppE phase terms (ppE_IMRPhenomD.cpp) --------------------------------------------------------------
MSOL_SEC = 4.925491025543575903411922162094833998e-6


def _ppe_case(model, nmod=2, b=(-7.0, -3.0), beta=(3e-5, -3e-3)):
    kw = dict(mass1=36., mass2=29., spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410., incl=.4, f_ref=20., phiRef=.3,
              tc=0.)
    gr = O.make_params(**kw)
    pp = O.make_params(ppe_nmod=nmod, ppe_b=list(b), ppe_beta=list(beta), **kw)
    mc = (36. * 29.) ** 0.6 / (65.) ** 0.2 * MSOL_SEC
    f = np.linspace(20., 0.2 / (65. * MSOL_SEC) * 0.999, 3001)

    def T(x):
        return sum(be * (np.pi * mc * x) ** (bb / 3.) for bb, be in zip(b[:nmod], beta[:nmod]))

    def dT(x):
        return sum(be * (bb / 3.) * (np.pi * mc * x) ** (bb / 3.) / x for bb, be in zip(b[:nmod], beta[:nmod]))
    return gr, pp, f, T, dT


def _linear_residual(f, d):
    A = np.vstack([np.ones_like(f), f]).T
    c, *_ = np.linalg.lstsq(A, d, rcond=None)
    return np.max(np.abs(d - A @ c))


def test_ppe_zero_coefficients_reduce_to_gr():
    for model in (O.PPE_IMRPHENOMD_INSPIRAL, O.PPE_IMRPHENOMD_IMR):
        gr, pp, f, T, dT = _ppe_case(model, beta=(0., 0.))
        a = O.fourier_waveform(f, gr, O.IMRPHENOMD)
        b = O.fourier_waveform(f, pp, model)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_ppe_imr_adds_the_term_in_every_region():
    """ppE_IMRPhenomD_IMR: the term enters all three phases, so the C1 connections are unchanged and the
    phase differs from GR by T(f) plus a line (the phase and time alignment shifts)"""
    gr, pp, f, T, dT = _ppe_case(O.PPE_IMRPHENOMD_IMR)
    g = O.lib()
    n = len(f)
    pg, pp_ = np.zeros(n), np.zeros(n)
    cross = np.zeros(n)
    g.gwo_fourier_phase(O._dp(f), n, O._dp(pg), O._dp(cross), O.IMRPHENOMD, O.ctypes.byref(gr))
    g.gwo_fourier_phase(O._dp(f), n, O._dp(pp_), O._dp(cross), O.PPE_IMRPHENOMD_IMR, O.ctypes.byref(pp))
    # fourier_phase returns -(Psi - tc (f - f_ref) - phic)
    d = -(pp_ - pg) - T(f)
    assert np.max(np.abs(T(f))) > 1.0  # the term is not negligible on this grid
    assert _linear_residual(f, d) < 1e-8 * np.max(np.abs(pg))
    # a wrong sign or exponent leaves a curved residual
    assert _linear_residual(f, -(pp_ - pg) + T(f)) > 1e-2
    assert _linear_residual(f, -(pp_ - pg) - 1.01 * T(f)) > 1e-4  # a 1 % coefficient error


def test_ppe_inspiral_term_is_carried_by_the_connections():
    """ppE_IMRPhenomD_Inspiral: the term enters the inspiral phase only; the C1 connection at f1 = 0.018/M
    carries its value and slope, so above f1 the phase differs from GR by the tangent of T at f1 (plus the
    common alignment line)"""
    gr, pp, f, T, dT = _ppe_case(O.PPE_IMRPHENOMD_INSPIRAL)
    g = O.lib()
    n = len(f)
    pg, pp_, cross = np.zeros(n), np.zeros(n), np.zeros(n)
    g.gwo_fourier_phase(O._dp(f), n, O._dp(pg), O._dp(cross), O.IMRPHENOMD, O.ctypes.byref(gr))
    g.gwo_fourier_phase(O._dp(f), n, O._dp(pp_), O._dp(cross), O.PPE_IMRPHENOMD_INSPIRAL, O.ctypes.byref(pp))
    f1 = 0.018 / (65. * MSOL_SEC)
    model_diff = np.where(f < f1, T(f), T(f1) + dT(f1) * (f - f1))
    d = -(pp_ - pg) - model_diff
    assert _linear_residual(f, d) < 1e-8 * np.max(np.abs(pg))
    # the IMR variant's form does not fit this one
    assert _linear_residual(f, -(pp_ - pg) - T(f)) > 1e-3


# ---- gIMR modifications (gIMRPhenomD.cpp) ----------------------------------------------------------------
def _gimr_phase(f, kw, model, **extra):
    p = O.make_params(**kw, **extra)
    n = len(f)
    ph, cross = np.zeros(n), np.zeros(n)
    O.lib().gwo_fourier_phase(O._dp(f), n, O._dp(ph), O._dp(cross), model, O.ctypes.byref(p))
    return ph


def _gimr_base():
    kw = dict(mass1=36., mass2=29., spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410., incl=.4, f_ref=20., phiRef=.3,
              tc=0.)
    f = np.linspace(20., 0.2 / (65. * MSOL_SEC) * 0.999, 3001)
    return kw, f


def test_gimr_zero_deltas_reduce_to_gr():
    kw, f = _gimr_base()
    a = O.fourier_waveform(f, O.make_params(**kw), O.IMRPHENOMD)
    b = O.fourier_waveform(f, O.make_params(**kw), O.GIMRPHENOMD)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("order", [-4, -3, -2, -1, 0, 1, 2, 3, 4, 7])
def test_gimr_phase_deltas_equal_the_ppe_mapping(order):
    """gIMRPhenomD::ppE_gIMR_mapping (gIMRPhenomD.cpp:24-52): delta_phi of PN order i is the ppE inspiral term
    b = i - 5, beta = (3/128) eta^(-i/5) phi_i delta (phi_i the GR coefficient, 1 for i < 0 and for i = 1,
    whose coefficient is set to delta) -- two independent code paths of the oracle must agree"""
    kw, f = _gimr_base()
    delta = 0.05 if order >= 0 else 2e-4
    dphi = [0.] * 14
    dphi[order + 4] = delta
    g = _gimr_phase(f, kw, O.GIMRPHENOMD, gimr_dphi=dphi)
    eta = 36. * 29. / 65. ** 2
    coeff = 1.0 if order < 0 or order == 1 else O.pn_phase_coefficients(O.make_params(**kw))[order]
    beta = 3. / 128. * eta ** (-order / 5.) * coeff * delta
    p = _gimr_phase(f, kw, O.PPE_IMRPHENOMD_INSPIRAL, ppe_nmod=1, ppe_b=[order - 5.], ppe_beta=[beta])
    gr = _gimr_phase(f, kw, O.IMRPHENOMD)
    assert np.max(np.abs(g - gr)) > 1e-2  # the modification is visible
    assert np.max(np.abs(g - p)) < 1e-9 * np.max(np.abs(gr)), np.max(np.abs(g - p))


def _gimr_f1p(kw):
    summ = np.zeros(32)
    O.lib().gwo_phenomd_setup_summary(O.ctypes.byref(O.make_params(**kw)), O.IMRPHENOMD, O._dp(summ))
    return summ[4]


def test_gimr_order5_is_a_constant_and_order6_the_ppe_term_of_b1():
    """phi order 5 scales the log-free 2.5PN coefficient c8: K delta c8 (pi M f)^0 is a constant in the
    inspiral, carried unchanged through the C1 connections, so the phase differs from GR by a line only.
    Order 6 scales the log-free 3PN coefficient c10: the ppE inspiral term b = 1 with
    beta = (3/128) eta^(-6/5) c10 delta (the ppE-gIMR mapping with phi_6 = c10)"""
    kw, f = _gimr_base()
    gr = _gimr_phase(f, kw, O.IMRPHENOMD)
    dphi = [0.] * 14
    dphi[5 + 4] = 0.1
    g5 = _gimr_phase(f, kw, O.GIMRPHENOMD, gimr_dphi=dphi)
    assert _linear_residual(f, g5 - gr) < 1e-9 * np.max(np.abs(gr))
    dphi = [0.] * 14
    dphi[6 + 4] = 0.05
    g6 = _gimr_phase(f, kw, O.GIMRPHENOMD, gimr_dphi=dphi)
    eta = 36. * 29. / 65. ** 2
    c10 = O.pn_phase_coefficients(O.make_params(**kw))[10]
    beta = 3. / 128. * eta ** (-6. / 5.) * c10 * 0.05
    p = _gimr_phase(f, kw, O.PPE_IMRPHENOMD_INSPIRAL, ppe_nmod=1, ppe_b=[1.], ppe_beta=[beta])
    assert np.max(np.abs(g6 - gr)) > 1e-2
    assert np.max(np.abs(g6 - p)) < 1e-9 * np.max(np.abs(gr)), np.max(np.abs(g6 - p))


@pytest.mark.parametrize("order", [8, 9])
def test_gimr_log_orders_connect_with_the_gr_log_derivative(order):
    """phi "orders" 8 and 9 scale the log parts c9, c11 of the 2.5PN and 3PN coefficients: the inspiral phase
    gains T(f) = K delta c9 ln x (order 8) or -K delta c11 x^(1/3) ln x (order 9), x = pi M f, K = 3/(128 eta).
    The reference's connection slope at f1 = 0.018/M differentiates the log terms with the GR coefficients
    (IMRPhenomD.cpp:1068-1084), so above f1 the phase carries T(f1) and only the part of T's slope that treats
    the log as a constant: 0 (order 8), -K delta c11 ln x1 x1^(-2/3) pi M / 3 (order 9).  Below f1 the phase
    is GR + T; the whole difference minus that model is one line, continuous at f1.  The self-consistent
    slope (T'(f1) itself) leaves a kink of K delta c9 / f1 (order 8) or K delta c11 x1^(1/3) / f1 (order 9)
    and fails."""
    kw, f = _gimr_base()
    delta = 0.1
    dphi = [0.] * 14
    dphi[order + 4] = delta
    g = _gimr_phase(f, kw, O.GIMRPHENOMD, gimr_dphi=dphi)
    gr = _gimr_phase(f, kw, O.IMRPHENOMD)
    M = 65. * MSOL_SEC
    eta = 36. * 29. / 65. ** 2
    K = 3. / (128. * eta)
    c = O.pn_phase_coefficients(O.make_params(**kw))
    f1 = _gimr_f1p(kw)
    x, x1 = np.pi * M * f, np.pi * M * f1
    if order == 8:
        T = lambda xx: K * delta * c[9] * np.log(xx)
        slope_ref, slope_true = 0.0, K * delta * c[9] / f1
    else:
        T = lambda xx: -K * delta * c[11] * xx ** (1. / 3.) * np.log(xx)
        slope_ref = -K * delta * c[11] * np.log(x1) * x1 ** (-2. / 3.) * np.pi * M / 3.
        slope_true = slope_ref - K * delta * c[11] * x1 ** (1. / 3.) / f1
    # fourier_phase returns -(Psi - tc (f - f_ref) - phic)
    below = f < f1
    ref_model = np.where(below, T(x), T(x1) + slope_ref * (f - f1))
    true_model = np.where(below, T(x), T(x1) + slope_true * (f - f1))
    scale = np.max(np.abs(gr))
    assert np.max(np.abs(T(x))) > 1e-2  # the term is visible
    assert _linear_residual(f, -(g - gr) - ref_model) < 1e-9 * scale
    assert _linear_residual(f, -(g - gr) - true_model) > 1e-4
    # the slope jump at f1 measured on the phase itself: one-sided cubic fits on either side
    h = 1e-3 * f1
    fl, fr = f1 - h * np.arange(1, 5), f1 + h * np.arange(1, 5)
    pl, pr = -_gimr_phase(fl, kw, O.GIMRPHENOMD, gimr_dphi=dphi), -_gimr_phase(fr, kw, O.GIMRPHENOMD, gimr_dphi=dphi)
    cl, cr = np.polyfit(fl - f1, pl, 3), np.polyfit(fr - f1, pr, 3)
    assert abs(cl[-1] - cr[-1]) < 1e-6 * (1 + abs(cl[-1]))  # C0 at f1
    jump = cr[-2] - cl[-2]
    expected = slope_ref - slope_true
    assert abs(jump - expected) < 1e-4 * abs(expected), (jump, expected)


@pytest.mark.parametrize("field,idx,region", [("gimr_dsigma", 0, "ins"), ("gimr_dsigma", 2, "ins"),
                                               ("gimr_dbeta", 0, "int"), ("gimr_dbeta", 1, "int"),
                                               ("gimr_dalpha", 0, "mr"), ("gimr_dalpha", 3, "mr")])
def test_gimr_fit_deltas_change_only_their_region(field, idx, region):
    """a calibration fit scales only its own region's phase; the C1 connections (and the time/phase
    alignment) add a line elsewhere, so the difference from GR is linear outside that region and not in it"""
    kw, f = _gimr_base()
    n = {"gimr_dsigma": 3, "gimr_dbeta": 2, "gimr_dalpha": 4}[field]
    d = [0.] * n
    d[idx] = 0.2
    diff = _gimr_phase(f, kw, O.GIMRPHENOMD, **{field: d}) - _gimr_phase(f, kw, O.IMRPHENOMD)
    M = 65. * MSOL_SEC
    summ = np.zeros(32)
    O.lib().gwo_phenomd_setup_summary(O.ctypes.byref(O.make_params(**kw)), O.IMRPHENOMD, O._dp(summ))
    f1p, f2p = summ[4], summ[5]
    regions = {"ins": f < f1p, "int": (f > f1p) & (f < f2p), "mr": f > f2p}
    for name, m in regions.items():
        if m.sum() < 20:
            continue
        res = _linear_residual(f[m], diff[m])
        if name == region:
            assert res > 1e-6, (name, res)
        else:
            assert res < 1e-9 * np.max(np.abs(diff)) + 1e-10, (name, res)


# ---- the same modifications on IMRPhenomPv2 (ppE_IMRPhenomP.cpp, gIMRPhenomP.cpp) ---------------------------
def _pv2_kw():
    return dict(mass1=36., mass2=29., spin1=[.3, .1, .2], spin2=[-.1, .2, .1], DL=410., incl=.4, f_ref=20.,
                phiRef=.3, tc=0.)


def _pv2_phase_shift(f, mod, model, **extra):
    """-arg(h_mod / h_Pv2) for both polarisations: the modifications change only the aligned-spin phase (and the
    t_c spline's slope), so the ratio is a pure phase, the same for h+ and hx"""
    kw = _pv2_kw()
    gp, gc = O.fourier_waveform(f, O.make_params(**kw), O.IMRPHENOMPV2)
    mp, mc = O.fourier_waveform(f, O.make_params(**kw, **extra), model)
    rp, rc = mp / gp, mc / gc
    assert np.max(np.abs(np.abs(rp) - 1)) < 1e-9 and np.max(np.abs(np.abs(rc) - 1)) < 1e-9
    assert np.max(np.abs(np.angle(rp / rc))) < 1e-9
    return -np.unwrap(np.angle(rp)), np.unwrap(np.angle(gp))


def test_pv2_variants_with_zero_terms_reduce_to_pv2():
    kw = _pv2_kw()
    f = np.linspace(20., 0.2 / (65. * MSOL_SEC) * 0.999, 2001)
    a = O.fourier_waveform(f, O.make_params(**kw), O.IMRPHENOMPV2)
    for model, extra in ((O.PPE_IMRPHENOMPV2_INSPIRAL, dict(ppe_nmod=2, ppe_b=[-7., -3.], ppe_beta=[0., 0.])),
                         (O.PPE_IMRPHENOMPV2_IMR, dict(ppe_nmod=1, ppe_b=[-7.], ppe_beta=[0.])),
                         (O.GIMRPHENOMPV2, {})):
        b = O.fourier_waveform(f, O.make_params(**kw, **extra), model)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # and the plain model ignores the modification fields
    c = O.fourier_waveform(f, O.make_params(**kw, ppe_nmod=1, ppe_b=[-7.], ppe_beta=[1e-4], gimr_dalpha=[.1] * 4),
                           O.IMRPHENOMPV2)
    assert np.array_equal(a[0], c[0])


def test_ppe_pv2_terms_and_connections():
    """ppE_IMRPhenomPv2_IMR: the aligned-spin phase gains T(f) in every region; ppE_IMRPhenomPv2_Inspiral: T(f)
    below f1 = 0.018/M and its tangent at f1 above (the C1 connections); both up to a line (alignment)"""
    _, _, f, T, dT = _ppe_case(O.PPE_IMRPHENOMD_IMR)
    extra = dict(ppe_nmod=2, ppe_b=[-7., -3.], ppe_beta=[3e-5, -3e-3])
    d, gph = _pv2_phase_shift(f, None, O.PPE_IMRPHENOMPV2_IMR, **extra)
    scale = np.max(np.abs(gph))
    assert np.max(np.abs(T(f))) > 1.0
    assert _linear_residual(f, d - T(f)) < 1e-8 * scale
    assert _linear_residual(f, d + T(f)) > 1e-2  # a wrong sign leaves a curve
    d, _ = _pv2_phase_shift(f, None, O.PPE_IMRPHENOMPV2_INSPIRAL, **extra)
    f1 = 0.018 / (65. * MSOL_SEC)
    assert _linear_residual(f, d - np.where(f < f1, T(f), T(f1) + dT(f1) * (f - f1))) < 1e-8 * scale
    assert _linear_residual(f, d - T(f)) > 1e-3  # the IMR variant's form does not fit


@pytest.mark.parametrize("order", [-3, 0, 2, 4, 7])
def test_gimr_pv2_phase_deltas_equal_the_ppe_mapping(order):
    """the gIMR <-> ppE mapping (gIMRPhenomD.cpp:24-52) holds on IMRPhenomPv2's aligned-spin phase too: the
    gIMRPhenomPv2 waveform with delta_phi of order i equals ppE_IMRPhenomPv2_Inspiral with b = i - 5,
    beta = (3/128) eta^(-i/5) phi_i delta"""
    kw = _pv2_kw()
    f = np.linspace(20., 0.2 / (65. * MSOL_SEC) * 0.999, 2001)
    delta = 0.05 if order >= 0 else 2e-4
    dphi = [0.] * 14
    dphi[order + 4] = delta
    eta = 36. * 29. / 65. ** 2
    coeff = 1.0 if order < 0 or order == 1 else O.pn_phase_coefficients(O.make_params(**kw))[order]
    beta = 3. / 128. * eta ** (-order / 5.) * coeff * delta
    dg, gph = _pv2_phase_shift(f, None, O.GIMRPHENOMPV2, gimr_dphi=dphi)
    dp, _ = _pv2_phase_shift(f, None, O.PPE_IMRPHENOMPV2_INSPIRAL, ppe_nmod=1, ppe_b=[order - 5.], ppe_beta=[beta])
    assert np.max(np.abs(dg)) > 1e-2
    assert np.max(np.abs(dg - dp)) < 1e-9 * np.max(np.abs(gph)), np.max(np.abs(dg - dp))


# ---- Fisher bases of the ppE methods (fisher.cpp:2083-2088, :2424-2429: betappe appended to the GR basis) ----
def _ppe_fisher_case(nmod=1, beta=(0.0,), b=(-7.0,)):
    kw = dict(mass1=36., mass2=29., spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410., incl=.4, f_ref=20., phiRef=.3,
              tc=0., RA=1.37, DEC=-1.26, psi=2.66, gmst=0.5)
    f = np.linspace(20., 0.2 / (65. * MSOL_SEC) * 0.999, 4001)
    psd = 1e-46 * (1 + (50. / f) ** 4 + (f / 300.) ** 2)
    p = O.make_params(ppe_nmod=nmod, ppe_b=list(b) + [0.] * (4 - nmod), ppe_beta=list(beta) + [0.] * (4 - nmod), **kw)
    mc = (36. * 29.) ** 0.6 / 65. ** 0.2 * MSOL_SEC
    return kw, f, psd, p, mc


def test_ppe_fisher_gr_block_at_beta_zero_is_the_gr_fisher():
    """at beta = 0 the ppE waveform and all its GR-parameter perturbations equal IMRPhenomD's, so the first 11 x 11
    block of the 12-dimensional ppE Fisher is the GR Fisher (an index or ordering mistake breaks it)"""
    kw, f, psd, p, mc = _ppe_fisher_case()
    g = O.fisher_numerical(f, psd, "IMRPhenomD", 0, 0, 11, O.make_params(**kw))
    for m in ("ppE_IMRPhenomD_IMR", "ppE_IMRPhenomD_Inspiral"):
        F = O.fisher_numerical(f, psd, m, 0, 0, 12, p)
        assert np.allclose(F[:11, :11], g, rtol=1e-12, atol=0), m
        assert F[11, 11] > 0


def test_ppe_fisher_beta_beta_element_closed_form():
    """d h / d beta = -i dphi/dbeta h with dphi/dbeta = T(f) - T(f_ref) for ppE_IMRPhenomD_IMR (the connections do
    not move; the phase alignment subtracts the value at f_ref; the Fisher runs without the time shift), and for
    ppE_IMRPhenomD_Inspiral the term below f1 = 0.018/M, its tangent at f1 above; so F_bb = 4 (|h|^2 dphi^2 / S)
    with the same Simpson rule"""
    kw, f, psd, p, mc = _ppe_fisher_case(beta=(2e-5,))
    h = O.detector_response(f, p, 0, O.PPE_IMRPHENOMD_IMR)
    T = lambda x: (np.pi * mc * x) ** (-7. / 3.)
    dT = lambda x: (-7. / 3.) * (np.pi * mc * x) ** (-7. / 3.) / x
    df = f[1] - f[0]
    F = O.fisher_numerical(f, psd, "ppE_IMRPhenomD_IMR", 0, 0, 12, p)
    d = T(f) - T(20.)
    ref = 4 * O.simpsons_sum(df, np.abs(h) ** 2 * d ** 2 / psd)
    assert abs(F[11, 11] - ref) <= 1e-6 * ref, (F[11, 11], ref)
    hi = O.detector_response(f, p, 0, O.PPE_IMRPHENOMD_INSPIRAL)
    f1 = 0.018 / (65. * MSOL_SEC)
    d = np.where(f < f1, T(f), T(f1) + dT(f1) * (f - f1)) - T(20.)
    ref = 4 * O.simpsons_sum(df, np.abs(hi) ** 2 * d ** 2 / psd)
    Fi = O.fisher_numerical(f, psd, "ppE_IMRPhenomD_Inspiral", 0, 0, 12, p)
    assert abs(Fi[11, 11] - ref) <= 1e-6 * ref, (Fi[11, 11], ref)
    # the other variant's closed form does not fit
    assert abs(Fi[11, 11] - F[11, 11]) > 1e-3 * ref


# ---- Fisher bases of the gIMR methods (fisher.cpp:2128-2158, :2473-2505: the listed deltas appended) ----------
def _gimr_fisher_case(codes, **tables):
    kw, f, psd, _, mc = _ppe_fisher_case()
    p = O.make_params(gimr_nlist=len(codes), gimr_list=codes, **tables, **kw)
    return kw, f, psd, p


# phi orders -2 and 2, sigma_3, beta_2, alpha_4 (codes: phi i -> i + 4, sigma i -> 12 + i, beta i -> 15 + i,
# alpha i -> 17 + i)
GIMR_CODES = [2, 6, 15, 17, 21]


def test_gimr_fisher_gr_block_at_zero_deltas_is_the_gr_fisher():
    """at zero deltas the gIMR waveform and its GR-parameter perturbations equal IMRPhenomD's: the first 11 x 11
    block of the gIMR Fisher is the GR Fisher, and every delta row has a positive diagonal"""
    kw, f, psd, p = _gimr_fisher_case(GIMR_CODES)
    g = O.fisher_numerical(f, psd, "IMRPhenomD", 0, 0, 11, O.make_params(**kw))
    F = O.fisher_numerical(f, psd, "gIMRPhenomD", 0, 0, 11 + len(GIMR_CODES), p)
    assert np.allclose(F[:11, :11], g, rtol=1e-12, atol=0)
    assert np.all(np.diag(F)[11:] > 0)


@pytest.mark.parametrize("order", [-2, 0, 2])
def test_gimr_fisher_phase_delta_is_the_scaled_ppe_beta(order):
    """delta_phi of PN order i is the ppE inspiral beta of b = i - 5 scaled by c = (3/128) eta^(-i/5) phi_i
    (gIMRPhenomD.cpp:24-52): at zero, the delta row of the gIMR Fisher is c times the beta row of the
    ppE_IMRPhenomD_Inspiral Fisher, and the delta-delta element c^2 times the beta-beta element. A code that
    addresses the wrong table entry, or a repack into the wrong order, breaks it."""
    kw, f, psd, p = _gimr_fisher_case([order + 4])
    F = O.fisher_numerical(f, psd, "gIMRPhenomD", 0, 0, 12, p)
    pp = O.make_params(ppe_nmod=1, ppe_b=[order - 5.], ppe_beta=[0.], **kw)
    Fp = O.fisher_numerical(f, psd, "ppE_IMRPhenomD_Inspiral", 0, 0, 12, pp)
    eta = 36. * 29. / 65. ** 2
    coeff = 1.0 if order < 0 else O.pn_phase_coefficients(O.make_params(**kw))[order]
    c = 3. / 128. * eta ** (-order / 5.) * coeff
    assert abs(F[11, 11] - c * c * Fp[11, 11]) <= 1e-6 * F[11, 11], (F[11, 11], c * c * Fp[11, 11])
    scale = np.sqrt(F[11, 11] * np.diag(F)[:11])
    assert np.max(np.abs(F[11, :11] - c * Fp[11, :11]) / scale) < 1e-6


@pytest.mark.parametrize("code", [15, 17, 21])
def test_gimr_fisher_fit_delta_against_a_direct_difference(code):
    """the delta-delta element of a fit delta (sigma_3, beta_2, alpha_4) against 4 sum |dh/ddelta|^2 / S with
    dh/ddelta a central difference of the waveform taken here, with the table entry the code names set directly
    (no time shift, as the Fisher's fresh record)"""
    kw, f, psd, p = _gimr_fisher_case([code])
    F = O.fisher_numerical(f, psd, "gIMRPhenomD", 0, 0, 12, p)
    field, idx, n = {15: ("gimr_dsigma", 1, 3), 17: ("gimr_dbeta", 0, 2), 21: ("gimr_dalpha", 2, 4)}[code]
    eps = 1e-6

    def resp(d):
        t = [0.] * n
        t[idx] = d
        return O.detector_response(f, O.make_params(**{**kw, field: t, "shift_time": 0}), 0, O.GIMRPHENOMD)

    dh = (resp(eps) - resp(-eps)) / (2 * eps)
    ref = 4 * O.simpsons_sum(f[1] - f[0], np.abs(dh) ** 2 / psd)
    assert abs(F[11, 11] - ref) <= 1e-5 * ref, (F[11, 11], ref)


def test_gimr_fisher_repeats_and_unread_indices_are_zero_rows():
    """the waveform reads the first list entry of an index (check_list_id), so a repeated entry and an index the
    model does not modify (code -1) are parameters the waveform does not depend on: zero rows and columns"""
    kw, f, psd, p = _gimr_fisher_case([6, -1, 6])
    F = O.fisher_numerical(f, psd, "gIMRPhenomD", 0, 0, 14, p)
    assert F[11, 11] > 0
    assert np.all(F[12] == 0) and np.all(F[13] == 0) and np.all(F[:, 12] == 0) and np.all(F[:, 13] == 0)


# ---- theories mapped onto ppE terms (ppE_utilities.cpp:158-360, waveform_generator.cpp:1379-1404) ----------
@pytest.mark.parametrize("cosmo,H0,Om", [(O.COSMO_PLANCK15, 67.74, 0.3075), (O.COSMO_PLANCK13, 67.77, 0.30712),
                                         (O.COSMO_WMAP9, 69.32, 0.2865), (O.COSMO_WMAP7, 70.4, 0.272),
                                         (O.COSMO_WMAP5, 70.2, 0.277)])
def test_redshift_fits_against_a_flat_lcdm_integral(cosmo, H0, Om):
    """Z_from_DL's fits (astropy cosmologies, include/gwat/D_Z_Config.h) against D_L(z) = (1+z) c/H0
    int_0^z dz'/E(z'), E^2 = Om (1+z)^3 + 1 - Om, inverted by root finding: within 3e-3 from 0.01 Mpc to
    100 Gpc (radiation and massive neutrinos, which the fits include, account for the rest); the next
    cosmology's H0 differs by more than that, so a wrong cosmology index or segment fails"""
    from scipy.integrate import quad
    from scipy.optimize import brentq
    ckm = 299792.458

    def dl(z):
        return (1 + z) * ckm / H0 * quad(lambda x: (Om * (1 + x) ** 3 + 1 - Om) ** -0.5, 0, z, epsrel=1e-12)[0]

    for d in (0.01, 0.5, 1.0, 10., 100., 400., 1000., 5000., 20000., 100000.):
        z = brentq(lambda x: dl(x) - d, 1e-12, 40)
        assert abs(O.z_from_dl(d, cosmo) - z) <= 3e-3 * z, (d, O.z_from_dl(d, cosmo), z)
    assert O.z_from_dl(1e6, cosmo) == -1  # past the table


def _theory(th, beta, **extra):
    kw = dict(mass1=36., mass2=29., spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410., incl=.4, f_ref=20., phiRef=.3,
              tc=0.)
    kw.update(extra)
    return O.make_params(ppe_theory=th, ppe_nmod=len(beta), ppe_beta=list(beta), **kw)


def test_theory_waveform_is_the_mapped_ppe_inspiral_waveform():
    """a theory's waveform is the ppE_IMRPhenomD_Inspiral / IMR waveform of its mapped terms (the mapping
    replaces the terms before the phase is built), bit for bit"""
    kw, f = _gimr_base()
    for th, beta, model in [(O.THEORY_DCS, [3e-22], O.PPE_IMRPHENOMD_INSPIRAL),
                            (O.THEORY_EDGB_GHOV3, [2e-22, 1.5], O.PPE_IMRPHENOMD_INSPIRAL),
                            (O.THEORY_TVG, [1e-11], O.PPE_IMRPHENOMPV2_INSPIRAL),
                            (O.THEORY_PPEALT, [1e-3, 2e-2], O.PPE_IMRPHENOMD_IMR)]:
        p = _theory(th, beta, ppe_b=[-7., -3.])
        b, be = O.theory_terms(p)
        q = O.make_params(**kw, ppe_nmod=len(b), ppe_b=list(b), ppe_beta=list(be))
        a = O.fourier_waveform(f, p, model)
        c = O.fourier_waveform(f, q, model)
        assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1]), th


def test_edgb_variants_against_each_other():
    """EdGB_GHOv2 with a unit generic coefficient repeats EdGB's beta at b = -5; GHOv3 multiplies it by the GR
    1PN phase coefficient 3715/756 + 55 eta / 9 (taken from the oracle's PN coefficients); in the small-spin
    limit the EdGB charges tend to 1 and EdGB's beta to -(EdGB_HO_LO's), both reducing to (1 - 4 eta) /
    eta^(18/5) (the reference's two conventions differ in sign)"""
    a = 2e-22
    bE, eE = O.theory_terms(_theory(O.THEORY_EDGB, [a]))
    b2, e2 = O.theory_terms(_theory(O.THEORY_EDGB_GHOV2, [a, 1.0]))
    b3, e3 = O.theory_terms(_theory(O.THEORY_EDGB_GHOV3, [a, 1.0]))
    assert list(b2) == [-7, -5] and list(bE) == [-7]
    assert e2[0] == eE[0] and abs(e2[1] - eE[0]) <= 1e-15 * abs(eE[0])
    phi2 = O.pn_phase_coefficients(O.make_params(mass1=36., mass2=29.))[2]
    assert abs(e3[1] - e2[1] * phi2) <= 1e-12 * abs(e3[1])
    b1, e1 = O.theory_terms(_theory(O.THEORY_EDGB_GHOV1, [a, 0.37]))
    assert e1[0] == eE[0] and e1[1] != 0
    small = dict(spin1=[0, 0, 1e-4], spin2=[0, 0, -3e-4])
    _, eS = O.theory_terms(_theory(O.THEORY_EDGB, [a], **small))
    _, eH = O.theory_terms(_theory(O.THEORY_EDGB_HO_LO, [a], **small))
    assert abs(eS[0] / eH[0] + 1) < 1e-6, eS[0] / eH[0]
    # the charges vanish below the reference's threshold: EdGB of non-spinning holes has no term
    assert O.theory_terms(_theory(O.THEORY_EDGB, [a], spin1=[0, 0, 0], spin2=[0, 0, 0]))[1][0] == 0


def test_dcs_equal_mass_equal_spin_closed_form():
    """equal masses and spins cancel the scalar-charge term (m1 s2 - m2 s1 = 0): beta = 16 pi alpha^2 / M_src^4
    (15075/114688) 4^(14/5) chi^2 (2 - 350/201) / 4, with M_src = M / (1 + z) at the fitted redshift; spins of
    both signs, and the redshift enters as (1 + z)^4"""
    for chi in (0.4, -0.7):
        p = _theory(O.THEORY_DCS, [1e-22], mass1=30., mass2=30., spin1=[0, 0, chi], spin2=[0, 0, chi], DL=900.)
        b, be = O.theory_terms(p)
        z = O.z_from_dl(900.)
        Ms = 60. * MSOL_SEC / (1 + z)
        ref = 16 * np.pi * 1e-22 / Ms ** 4 * (15075. / 114688.) * 4 ** 2.8 * chi ** 2 * (2 - 350. / 201.) / 4
        assert list(b) == [-1] and abs(be[0] - ref) <= 1e-12 * abs(ref), (be[0], ref)


def test_pnseries_is_ppealt_with_scaled_coefficients():
    """PNSeries_beta's terms beyond the first carry beta_0 beta_i where ppEAlt_beta carries beta_i; both convert
    the M f basis to Mc f by (M/Mc)^(b/3) (two code paths of the oracle)"""
    b = [-7., -5., -3.]
    _, s = O.theory_terms(_theory(O.THEORY_PNSERIES, [2e-3, 0.5, -4.], ppe_b=b))
    _, t = O.theory_terms(_theory(O.THEORY_PPEALT, [2e-3, 2e-3 * 0.5, 2e-3 * -4.], ppe_b=b))
    assert np.allclose(s, t, rtol=1e-15, atol=0)
    eta = 36. * 29. / 65. ** 2
    assert abs(t[0] - 2e-3 * eta ** (-0.6 * -7. / 3.)) <= 1e-12 * abs(t[0])  # M / Mc = eta^(-3/5)


def test_theory_refusals():
    """GHO variants read two theory parameters; a theory outside the list, an unknown cosmology: no terms"""
    assert O.theory_terms(_theory(O.THEORY_EDGB_GHOV1, [1e-22])) is None
    assert O.theory_terms(_theory(O.THEORY_MODDISPERSION + 1, [1e-22])) is None
    assert O.theory_terms(_theory(O.THEORY_DCS, [1e-22], cosmology=9)) is None


@pytest.mark.parametrize("alpha", [0., 0.5, 1., 2., 2.5, 4.])
def test_modified_dispersion_distance_against_its_integral(alpha):
    """DL_from_Z_MD's fits (Planck15) against D_alpha = (1+z)^(1-alpha) c/H0 int_0^z (1+z')^(alpha-2)/E(z') dz'
    (Mirshekari, Yunes, Will 2012, arXiv:1110.2720) with flat LCDM, within 3e-3 for z = 1e-4 .. 10; a neighbouring
    alpha's fit differs by far more at z = 2"""
    from scipy.integrate import quad
    L = O.lib()
    L.gwo_md_distance.restype = O.ctypes.c_double
    L.gwo_md_distance.argtypes = [O.ctypes.c_double, O.ctypes.c_double]
    H0, Om, ckm = 67.74, 0.3075, 299792.458

    def D(z, a):
        I = quad(lambda x: (1 + x) ** (a - 2) / np.sqrt(Om * (1 + x) ** 3 + 1 - Om), 0, z, epsrel=1e-12)[0]
        return (1 + z) ** (1 - a) * ckm / H0 * I

    for z in (1e-4, 0.01, 0.087, 0.5, 2.0, 10.):
        assert abs(L.gwo_md_distance(z, alpha) / D(z, alpha) - 1) < 3e-3, z
    other = alpha + 0.5 if alpha < 4 else alpha - 0.5
    assert abs(L.gwo_md_distance(2.0, other) / D(2.0, alpha) - 1) > 0.1
    assert L.gwo_md_distance(0.5, 1.25) == -1  # no fit for that power


def test_moddispersion_terms():
    """ModDispersion keeps the given exponent b = 3 alpha - 3 and maps A_alpha to beta (ModDispersion_beta); the
    beta is linear in A_alpha; powers without a distance fit are refused"""
    p = _theory(O.THEORY_MODDISPERSION, [3e-15], ppe_b=[4.5])
    b, be = O.theory_terms(p)
    _, be2 = O.theory_terms(_theory(O.THEORY_MODDISPERSION, [6e-15], ppe_b=[4.5]))
    assert list(b) == [4.5] and be[0] != 0 and abs(be2[0] / be[0] - 2) < 1e-14
    assert O.theory_terms(_theory(O.THEORY_MODDISPERSION, [3e-15], ppe_b=[1.])) is None


@pytest.mark.parametrize("th,val,b", [(O.THEORY_NONCOMM, 0.05, -1.), (O.THEORY_DIPRAD, 1e-4, -7.)])
def test_theory_fisher_parameter_element(th, val, b):
    """a theory parameter whose beta is linear in it (beta = c x): its Fisher diagonal is c^2 times the ppE beta-beta
    element at the same waveform (the theory map and redshift travel into every perturbed record), while the
    GR block differs from the ppE one (beta also moves with the masses)"""
    kw, f, psd, _, mc = _ppe_fisher_case()
    p = O.make_params(ppe_theory=th, ppe_nmod=1, ppe_beta=[val], **kw)
    bb, be = O.theory_terms(p)
    c = be[0] / val
    q = O.make_params(ppe_nmod=1, ppe_b=[b], ppe_beta=[be[0]], **kw)
    F = O.fisher_numerical(f, psd, "ppE_IMRPhenomD_Inspiral", 0, 0, 12, p, epsilon=1e-6)
    # the same step in beta (c eps) for the beta-beta element, so both difference quotients match
    Fp = O.fisher_numerical(f, psd, "ppE_IMRPhenomD_Inspiral", 0, 0, 12, q, epsilon=1e-6 * abs(c))
    Fg = O.fisher_numerical(f, psd, "ppE_IMRPhenomD_Inspiral", 0, 0, 12, q, epsilon=1e-6)
    assert list(bb) == [b]
    assert abs(F[11, 11] - c * c * Fp[11, 11]) <= 1e-5 * F[11, 11], (F[11, 11], c * c * Fp[11, 11])
    assert np.max(np.abs(F[:11, :11] - Fg[:11, :11]) / np.sqrt(np.outer(np.diag(Fg)[:11], np.diag(Fg)[:11]))) > 1e-8


# ---- IMRPhenomPv2 precession pieces (IMRPhenomP.cpp) --------------------------------------------------------
def _wigner_d2(beta):
    """d^2_{m'm}(beta) = <2 m'| exp(-i beta J_y) |2 m> from the spin-2 matrices (a matrix exponential, no
    closed forms), indexed [m' + 2, m + 2]"""
    from scipy.linalg import expm
    ms = np.arange(-2, 3)
    jp = np.zeros((5, 5))
    for a, m in enumerate(ms[:-1]):
        jp[a + 1, a] = np.sqrt(6 - m * (m + 1))  # J+ |2 m> = sqrt(j(j+1) - m(m+1)) |2 m+1>
    jy = (jp - jp.T) / 2j
    return expm(-1j * beta * jy).real


def _pv2_cases():
    return [dict(mass1=36., mass2=29., spin1=[.3, .1, .2], spin2=[-.1, .2, .1], DL=410., incl=.4, f_ref=20.,
                 phiRef=.3),
            dict(mass1=50., mass2=8., spin1=[.7, -.5, -.3], spin2=[.2, .0, .4], DL=900., incl=2.2, f_ref=20.,
                 phiRef=1.9),
            dict(mass1=12., mass2=10., spin1=[-.05, .9, .1], spin2=[.6, .6, -.2], DL=200., incl=1.4, f_ref=30.,
                 phiRef=4.0)]


@pytest.mark.parametrize("case", range(3))
def test_pv2_wigner_d_against_the_spin2_rotation_matrix(case):
    """IMRPhenomP.cpp:664-693: d2[m+2] = d^2_{m,2}(beta), dm2[m+2] = d^2_{m,-2}(beta) with cos beta from the
    2PN angular momentum; checked against exp(-i beta J_y) of spin 2, and unitarity sum_m |d^2_{m,2}|^2 = 1"""
    p = O.make_params(**_pv2_cases()[case])
    for f in (20., 45., 120., 300.):
        pc = O.pv2_pieces(p, f)
        beta = np.arccos(pc["cos_beta"])
        assert beta > 1e-3  # precessing: the test is not the trivial beta = 0
        D = _wigner_d2(beta)
        np.testing.assert_allclose(pc["d2"], D[:, 4], rtol=0, atol=1e-14)
        np.testing.assert_allclose(pc["dm2"], D[:, 0], rtol=0, atol=1e-14)
        assert abs(np.sum(pc["d2"] ** 2) - 1) < 1e-14 and abs(np.sum(pc["dm2"] ** 2) - 1) < 1e-14


@pytest.mark.parametrize("case", range(3))
def test_pv2_spin_weighted_harmonics(case):
    """util.cpp:1953-2000 at s = -2, l = 2, phi = 0: -2Y_2m(theta) = sqrt(5 / 4 pi) d^2_{m,2}(theta) (Goldberg
    et al.), and each is normalised on the sphere"""
    from scipy.integrate import quad
    p = O.make_params(**_pv2_cases()[case])
    th = O.pv2_setup_summary(p)["thetaJN"]
    y = O.pv2_pieces(p, 50.)["ylm"]
    np.testing.assert_allclose(y, np.sqrt(5 / (4 * np.pi)) * _wigner_d2(th)[:, 4], rtol=0, atol=1e-14)
    for m in range(-2, 3):
        norm = quad(lambda t: 2 * np.pi * np.sin(t) * (5 / (4 * np.pi)) * _wigner_d2(t)[m + 2, 4] ** 2, 0, np.pi)[0]
        assert abs(norm - 1) < 1e-10


def test_pv2_reduced_transform_equals_the_full_one():
    """PhenomPv2_Param_Transform_reduced (IMRPhenomP.cpp:821-938) takes chi_p, phi_p; with in-plane spin on the
    larger black hole only, chi_p = |chi_perp| and phi_p its azimuth give the full transform (:1081-1211)"""
    f = np.arange(20, 1024.01, 0.25)
    for s1 in ([.3, -.2, .2], [-.6, .1, -.4], [.0, .8, .1]):
        kw = dict(mass1=36., mass2=29., spin1=s1, spin2=[0, 0, .1], DL=410., incl=.4, f_ref=20., phiRef=.3)
        a = O.fourier_waveform(f, O.make_params(**kw), O.IMRPHENOMPV2)
        chip, phip = np.hypot(s1[0], s1[1]), np.arctan2(s1[1], s1[0])
        assert abs(O.pv2_setup_summary(O.make_params(**kw))["chip"] - chip) < 1e-15
        b = O.fourier_waveform(f, O.make_params(**kw, chip=chip, phip=phip), O.IMRPHENOMPV2)
        for x, y in zip(a, b):
            assert np.max(np.abs(x - y)) < 1e-12 * np.max(np.abs(x))
        # a wrong azimuth moves the waveform
        c = O.fourier_waveform(f, O.make_params(**kw, chip=chip, phip=phip + 0.3), O.IMRPHENOMPV2)
        assert np.max(np.abs(c[0] - a[0])) > 1e-3 * np.max(np.abs(a[0]))


@pytest.mark.parametrize("case", range(3))
def test_pv2_zeta_rotation_is_a_polarisation_angle_shift(case):
    """waveform_generator.cpp:257-266 rotates (h+, hx) by 2 zeta: the norm |h+|^2 + |hx|^2 is unchanged, and
    the detector response of the rotated waveform at psi is that of the unrotated one at psi - zeta (the
    antenna pattern's own polarisation-angle rotation); a reflection, a rotation by zeta or by -2 zeta fails"""
    f = np.arange(20, 1024.01, 0.25)
    kw = dict(_pv2_cases()[case], RA=1.0, DEC=0.3, psi=0.2, gmst=0.5)
    p = O.make_params(**kw)
    z = O.pv2_setup_summary(p)["zeta"]
    assert abs(np.sin(2 * z)) > 0.05
    hp, hc = O.fourier_waveform(f, p, O.IMRPHENOMPV2)
    up, uc = O.pv2_waveform_unrotated(f, p)
    n1, n0 = np.abs(hp) ** 2 + np.abs(hc) ** 2, np.abs(up) ** 2 + np.abs(uc) ** 2
    nz = n0 > 0
    assert np.max(np.abs(n1[nz] / n0[nz] - 1)) < 1e-12
    for det in (0, 1, 2):
        Gp, Gx = O.antenna(det, 1.0, 0.3, 0.2, 0.5)
        Fp, Fx = O.antenna(det, 1.0, 0.3, 0.2 - z, 0.5)
        a, b = Gp * hp + Gx * hc, Fp * up + Fx * uc
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))
        Fp2, Fx2 = O.antenna(det, 1.0, 0.3, 0.2 + z, 0.5)
        assert np.max(np.abs(a - (Fp2 * up + Fx2 * uc))) > 1e-2 * np.max(np.abs(a))


@pytest.mark.parametrize("case", range(3))
def test_pv2_time_correction_is_the_phase_slope_at_ringdown(case):
    """calculate_time_shift (IMRPhenomP.cpp:550-606): t_corr = d(-Psi)/df / 2 pi at f_RD, from a 10-knot natural
    spline on [0.8, 1.2] f_RD of the aligned-spin phase; it equals the analytic merger-ringdown slope and a
    central difference of the phase to the spline's accuracy; a lost sign or 2 pi fails"""
    p = O.make_params(**_pv2_cases()[case])
    s = O.pv2_setup_summary(p)
    fr = s["fRD"]
    h = 1e-3 * fr
    fd = -(O.pv2_aligned_phase(p, fr + h) - O.pv2_aligned_phase(p, fr - h)) / (2 * h) / (2 * np.pi)
    an = -s["dphase_mr_at_fRD"] / (2 * np.pi)
    assert abs(fd / an - 1) < 1e-5
    assert abs(s["t_corr"] / an - 1) < 1e-3, (s["t_corr"], an)


@pytest.mark.parametrize("m1,m2,chil", [(36., 29., 0.), (50., 8., 0.), (12., 10., 0.3), (30., 30., 0.), (50., 8., 0.4)])
def test_pv2_euler_angles_leading_order(m1, m2, chil):
    """IMRPhenomP.cpp:64-144 (the NNLO single-spin Euler angles, spin on the larger body b, s the other): as
    omega = pi M f -> 0 the angles diverge as c1 / omega, and c1 is fixed by leading-order physics.  L precesses
    about J at Omega_p = (2 + 3 m_s / (2 m_b)) L / r^3 = (2 + 3 m_s / (2 m_b)) eta v^5 / M (spin-orbit precession,
    Apostolatos et al. 1994, Kidder 1995) while the orbit chirps at d(M omega)/dt = (96/5) eta v^11 / M (v^3 = M
    omega), so d alpha / d(M omega) = (5/96) (2 + 3 m_s / (2 m_b)) (M omega)^-2 and c1 = -(5/96)(2 + 3 m_s / (2 m_b));
    epsilon' = alpha' cos(beta) with beta -> 0 at leading order has the same c1.  The omega^-2/3 coefficient c2
    follows from the same precession rate with J = L + S_L instead of L (L = eta / v, S_L = chi_L m_b^2): the
    relative correction chi_L m_b^2 v / eta integrates to c2 = (3/2) c1 chi_L m_b^2 / eta, for both angles (cos
    beta differs from 1 only at relative order v^2).  The coefficients are read off the oracle's angles at five tiny
    frequencies by a fit over 1/omega, omega^-2/3, omega^-1/3 and a constant (the f_ref offsets); the next terms are
    suppressed by further powers of omega^1/3.  A wrong mass ratio, a swapped spinning body, a lost factor or sign
    in either coefficient fails; c3 (omega^-1/3), c4 (log omega) and c5 (omega^1/3) stay unpinned."""
    kw = dict(mass1=m1, mass2=m2, spin1=[.4, .2, chil], spin2=[0., 0., 0.], DL=410., incl=.4, f_ref=20., phiRef=.3)
    p = O.make_params(**kw)
    M = 0.2 / O.fcut(p)
    fs = np.array([1e-6, 2e-6, 4e-6, 8e-6, 1.6e-5])
    om = np.pi * M * fs
    A = np.stack([1 / om, om ** (-2 / 3), om ** (-1 / 3), np.ones_like(om)], axis=1)
    scale = np.abs(A).max(axis=0)
    mb, ms = max(m1, m2) / (m1 + m2), min(m1, m2) / (m1 + m2)
    eta = mb * ms
    chi_l = O.pv2_setup_summary(p)["chil"]
    assert abs(chi_l - chil) < 1e-12  # the in-plane spin sits on the larger body only
    c1 = -(5 / 96) * (2 + 3 * ms / (2 * mb))
    c2 = 1.5 * c1 * chi_l * mb ** 2 / eta
    for key in ("alpha", "epsilon"):
        y = np.array([O.pv2_pieces(p, f)[key] for f in fs])
        coef = np.linalg.lstsq(A / scale, y, rcond=None)[0] / scale
        assert abs(coef[0] / c1 - 1) < 1e-6, (key, coef[0], c1)
        assert abs(coef[1] - c2) < 2e-4 + 1e-3 * abs(c2), (key, coef[1], c2)
    # the same fit tells a planted mistake apart: the mass ratio inverted in the leading coefficient
    wrong = -(5 / 96) * (2 + 3 * mb / (2 * ms))
    assert m1 == m2 or abs(wrong / c1 - 1) > 1e-2


@pytest.mark.parametrize("m1,m2,chil", [(36., 29., 0.2), (50., 8., 0.4), (12., 10., -0.3)])
def test_pv2_euler_angles_inplane_spin_term(m1, m2, chil):
    """The in-plane spin's part of the Euler angles (IMRPhenomP.cpp:64-144).  The precession rate of L about J
    scales with |J| = sqrt((L + S_L)^2 + S_p^2) = L + S_L + S_p^2 / (2 L) - S_p^2 S_L / (2 L^2) + ... (L = eta / v,
    S_p = chi_p m_b^2, S_L = chi_L m_b^2): relative corrections chi_p^2 m_b^4 v^2 / (2 eta^2) and
    -chi_p^2 chi_L m_b^6 v^3 / (2 eta^3), which integrate to an omega^-1/3 (c3) term (3/2) c1 chi_p^2 m_b^4 / eta^2
    and a log omega (c4) term c1 chi_p^2 chi_L m_b^6 / (2 eta^3) in alpha.  In epsilon' = alpha' cos(beta) =
    alpha' (L + S_L) / |J| they cancel.  The next term of the expansion, -S_p^4 / (8 L^3), gives the omega^1/3 (c5)
    coefficient's chi_p^4 part (3/8) c1 m_b^8 / eta^4.  Checked from the fitted coefficients at chi_p = 0.2, 0.5 and
    0.8: the changes of c3 and c4, and the chi_p^4 term of a quadratic in chi_p^2 through c5 (the spin-independent
    parts drop out)."""
    def fit(p, key):
        M = 0.2 / O.fcut(p)
        fs = np.geomspace(1e-6, 2e-3, 14)
        om = np.pi * M * fs
        A = np.stack([1 / om, om ** (-2 / 3), om ** (-1 / 3), np.log(om), np.ones_like(om), om ** (1 / 3)], axis=1)
        sc = np.abs(A).max(axis=0)
        y = np.array([O.pv2_pieces(p, f)[key] for f in fs])
        return np.linalg.lstsq(A / sc, y, rcond=None)[0] / sc
    co = {}
    for chip in (0.2, 0.5, 0.8):
        p = O.make_params(mass1=m1, mass2=m2, spin1=[chip, 0., chil], spin2=[0., 0., 0.], DL=410., incl=.4,
                          f_ref=20., phiRef=.3)
        assert abs(O.pv2_setup_summary(p)["chip"] - chip) < 1e-12
        co[chip] = (fit(p, "alpha"), fit(p, "epsilon"))
    mb, ms = max(m1, m2) / (m1 + m2), min(m1, m2) / (m1 + m2)
    eta = mb * ms
    c1 = -(5 / 96) * (2 + 3 * ms / (2 * mb))
    dchip2 = 0.8 ** 2 - 0.2 ** 2
    want3 = 1.5 * c1 * dchip2 * mb ** 4 / eta ** 2
    want4 = c1 * dchip2 * chil * mb ** 6 / (2 * eta ** 3)
    da, de = co[0.8][0] - co[0.2][0], co[0.8][1] - co[0.2][1]
    assert abs(da[2] / want3 - 1) < 1e-5, (da[2], want3)
    assert abs(da[3] / want4 - 1) < 1e-3, (da[3], want4)
    assert np.all(np.abs(de[:4]) < 1e-6 * abs(want3)), de
    quartic = np.polyfit(np.array([0.2, 0.5, 0.8]) ** 2, [co[c][0][5] for c in (0.2, 0.5, 0.8)], 2)[0]
    want5 = 3 / 8 * c1 * mb ** 8 / eta ** 4
    assert abs(quartic / want5 - 1) < 1e-2, (quartic, want5)


@pytest.mark.parametrize("m1,m2", [(36., 29.), (50., 8.), (12., 10.), (30., 30.), (40., 20.)])
def test_pv2_euler_angles_1pn_and_tail_terms(m1, m2):
    """The Euler angles' omega^-1/3 and log omega coefficients without aligned spin (IMRPhenomP.cpp:64-144).  With
    the spinning body's precession rate at next-to-leading order, Omega = (v^5 / M) [A0 + A2 v^2], A0 = 3/4 + eta/2
    - 3 D / 4, A2 = 9/16 + 5 eta / 4 - eta^2 / 24 + D (-9/16 + 5 eta / 8) (D = m_b - m_s; the NLO spin-orbit
    precession frequency of the post-Newtonian literature, e.g. Blanchet, Living Rev. Relativ. 17, 2 (2014)), over
    the chirp d(M omega)/dt = (96/5) eta v^11 [1 + b2 v^2 + b3 v^3] / M with b2 = -(743/336 + 11 eta / 4) and the
    tail term b3 = 4 pi: the omega^-1/3 coefficient is 3 c1 (A2 / A0 - b2) (plus the in-plane spin's
    (3/2) c1 chi_p^2 m_b^4 / eta^2 in alpha) and the log omega coefficient 4 pi c1.  The aligned-spin parts of the log
    omega and omega^1/3 coefficients stay unpinned."""
    kw = dict(mass1=m1, mass2=m2, spin1=[.5, 0., 0.], spin2=[0., 0., 0.], DL=410., incl=.4, f_ref=20., phiRef=.3)
    p = O.make_params(**kw)
    chip = O.pv2_setup_summary(p)["chip"]
    M = 0.2 / O.fcut(p)
    fs = np.geomspace(1e-6, 2e-4, 10)
    om = np.pi * M * fs
    A = np.stack([1 / om, om ** (-2 / 3), om ** (-1 / 3), np.log(om), np.ones_like(om), om ** (1 / 3)], axis=1)
    sc = np.abs(A).max(axis=0)
    mb, ms = max(m1, m2) / (m1 + m2), min(m1, m2) / (m1 + m2)
    eta, D = mb * ms, max(m1, m2) / (m1 + m2) - min(m1, m2) / (m1 + m2)
    c1 = -(5 / 96) * (2 + 3 * ms / (2 * mb))
    a0 = 3 / 4 + eta / 2 - 3 * D / 4
    a2 = 9 / 16 + 5 * eta / 4 - eta ** 2 / 24 + D * (-9 / 16 + 5 * eta / 8)
    b2 = -(743 / 336 + 11 * eta / 4)
    assert abs(a0 / eta - (2 + 3 * ms / (2 * mb))) < 1e-14  # the leading rate of the precession test above
    c3 = 3 * c1 * (a2 / a0 - b2)
    for key, extra in (("alpha", 1.5 * c1 * chip ** 2 * mb ** 4 / eta ** 2), ("epsilon", 0.0)):
        y = np.array([O.pv2_pieces(p, f)[key] for f in fs])
        coef = np.linalg.lstsq(A / sc, y, rcond=None)[0] / sc
        assert abs(coef[2] / (c3 + extra) - 1) < 1e-5, (key, coef[2], c3 + extra)
        assert abs(coef[3] / (4 * np.pi * c1) - 1) < 1e-3, (key, coef[3], 4 * np.pi * c1)
    # a dropped 11 eta / 4 in the chirp's 1PN term moves c3 by more than the fit's tolerance
    assert abs(3 * c1 * (a2 / a0 + 743 / 336) / c3 - 1) > 1e-3


# ---- the order-4 Fisher stencil (fisher.cpp:403-453) -------------------------------------------------------
def test_order4_stencil_is_the_five_point_rule_and_its_tc_shift_mix():
    """Order 4: (-h(+2e) + 8 h(+e) - 8 h(-e) + h(-2e)) / 12 e.  With the detector equal to the reference one it
    is the textbook five-point rule of the response.  Otherwise the +-e responses carry the time-of-flight
    factor e^{-2 pi i dt f} while the +-2e ones shift t_c by -dt instead, which (h depends on t_c through
    e^{2 pi i t_c (f - f_ref)}) is the same factor times e^{2 pi i dt f_ref}: the derivative differs from the
    consistent five-point rule by (e^{i k} - 1)(h(-2e) - h(+2e)) / 12 e with k = 2 pi dt f_ref per point."""
    f = np.arange(20., 512.01, 0.5)
    kw = dict(mass1=36., mass2=29., spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410., incl=.4, f_ref=20., phiRef=.3,
              RA=1.37, DEC=-1.26, psi=2.66, gmst=0.7, tc=0.01, shift_time=0)
    p = O.make_params(**kw)
    dim, eps, method = 11, 1e-5, "IMRPhenomD"
    v, lf = O.unpack_parameters(p, method, dim)

    def resp(vv, det, ref):
        q = O.repack_parameters(vv, method, dim, p)
        dt = O.dtoa(q.RA, q.DEC, q.gmst, ref, det)
        r = O.detector_response(f, q, det)
        q.tc -= dt
        return r * np.exp(-2j * np.pi * dt * f), O.detector_response(f, q, det), dt

    # the t_c shift equals the time-of-flight factor times e^{2 pi i dt f_ref}
    a, b, dt = resp(v, 1, 0)
    assert abs(dt) > 1e-3
    assert np.max(np.abs(b - a * np.exp(2j * np.pi * dt * 20.))) < 1e-12 * np.max(np.abs(a))
    for det, ref in ((0, 0), (1, 0), (2, 1)):
        D = O.calculate_derivatives(f, method, det, ref, dim, p, order=4, epsilon=eps)
        for i in range(dim):
            e = np.zeros(dim)
            e[i] = eps
            (c_p, _, _), (c_m, _, _) = resp(v + e, det, ref), resp(v - e, det, ref)
            (c_pp, s_pp, _), (c_mm, s_mm, _) = resp(v + 2 * e, det, ref), resp(v - 2 * e, det, ref)
            consistent = (-c_pp + 8. * c_p - 8. * c_m + c_mm) / (12. * eps)
            mix = ((c_mm - s_mm) - (c_pp - s_pp)) / (12. * eps) if det != ref else 0
            scale = v[i] if lf[i] else 1.0
            want = (consistent - mix) * scale
            # rounding of the four responses (a few ulp each, time-of-flight factors from numpy's exp here and
            # libm's sin/cos in the oracle) divided by 12 e, plus a relative floor
            noise = 18 * 4.4e-16 * np.max(np.abs(c_p)) / (12. * eps) * abs(scale)
            assert np.max(np.abs(D[i] - want)) < noise + 1e-12 * np.max(np.abs(want)), (det, i)
            if det != ref and i == 5:  # the mix is not negligible (t_c derivative)
                assert np.max(np.abs(mix)) > 1e-3 * np.max(np.abs(consistent))


def test_relative_step_is_the_absolute_step_of_each_coordinate():
    """epsilon < 0 asks for the step |epsilon| * |v_i| per basis coordinate (|epsilon| where v_i is 0): each
    derivative row equals the absolute-step row at that coordinate's step, and the Fisher closed forms of the
    full basis hold with it (F(lnDL,lnDL) = (h|h), F(phi,phi) = 4 (h|h), F(tc,tc) at tc = 0)."""
    f = np.arange(20, 1024.01, 0.25)
    x = f / 215.
    psd = 1e-49 * (x ** -4.14 - 5 * x ** -2 + 111 * (1 - x ** 2 + x ** 4 / 2) / (1 + x ** 2 / 2))
    rel = 1e-7
    for method, dim, sky in (("IMRPhenomD", 7, True), ("IMRPhenomD", 11, False), ("MCMC_IMRPhenomD", 11, False)):
        p = O.make_params(mass1=36, mass2=29, spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410, incl=2.9, RA=1.37,
                          DEC=-1.26, psi=2.66, f_ref=20, tc=0., phiRef=0., sky_average=sky)
        v, _ = O.unpack_parameters(p, method, dim)
        assert np.any(v == 0) and np.any(np.abs(v) > 1)
        D = O.calculate_derivatives(f[:801], method, 1, 0, dim, p, order=2, epsilon=-rel)
        for i in range(dim):
            h = rel * abs(v[i]) if v[i] != 0 else rel
            A = O.calculate_derivatives(f[:801], method, 1, 0, dim, p, order=2, epsilon=h)
            assert np.array_equal(D[i], A[i]), (method, i)
    p = O.make_params(mass1=36, mass2=29, spin1=[0, 0, .3], spin2=[0, 0, -.1], DL=410, incl=2.9, RA=1.37,
                      DEC=-1.26, psi=2.66, f_ref=20, tc=0.)
    F = O.fisher_numerical(f, psd, "IMRPhenomD", 0, 0, 11, p, epsilon=-rel)
    r = O.detector_response(f, p, 0)
    hh = 4 * O.simpsons_sum(f[1] - f[0], np.abs(r) ** 2 / psd)
    assert abs(F[6, 6] / hh - 1) < 1e-5
    assert abs(F[4, 4] / (4 * hh) - 1) < 1e-5
    ftt = 4 * O.simpsons_sum(f[1] - f[0], (2 * np.pi * (f - 20.)) ** 2 * np.abs(r) ** 2 / psd)
    assert abs(F[5, 5] / ftt - 1) < 1e-4
    # a planted error: a step 10x the coordinate's in one row would move that row's derivative
    D10 = O.calculate_derivatives(f[:801], "IMRPhenomD", 1, 0, 11, p, order=2, epsilon=10 * rel * abs(v[0]))
    D1 = O.calculate_derivatives(f[:801], "IMRPhenomD", 1, 0, 11, p, order=2, epsilon=-rel)
    assert not np.array_equal(D10[0], D1[0])


# ---- Gauss-Legendre and log10-f integration rules (mcmc_gw.cpp:821-856, fisher.cpp:2736-2745) ------------------
def _gl(x1, x2, n):
    """Gauss-Legendre nodes and weights on [x1, x2] (numpy's leggauss: the same rule gauleg builds,
    ortho_basis.cpp:14)"""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x2 - x1) * x + 0.5 * (x2 + x1), 0.5 * (x2 - x1) * w


def test_gauss_legendre_inner_products_against_closed_forms():
    """the GAUSSLEG branch of Log_Likelihood_internal: (a|b) = 4 sum w_i Re(a_i b_i*) / S_i, times f_i ln 10 on a
    log10 grid (df = f ln 10 d log10 f).  With power laws h = f^a e^{i phi}, d = f^b e^{i phi}, S = f^c the inner
    products are 4 int f^(2a - c) df and 4 int f^(a + b - c) df in closed form: exact for polynomial integrands on
    a linear grid, to rounding on a log10 grid (smooth in log10 f); a dropped Jacobian or weight fails."""
    f1, f2 = 20., 1024.
    prim = lambda k: 4 * (f2 ** (k + 1) - f1 ** (k + 1)) / (k + 1)
    # linear grid, polynomial integrands (a = 1.5, b = 0.5, c = 0 -> f^3 and f^2)
    f, w = _gl(f1, f2, 32)
    ph = np.exp(1j * 0.37 * f)
    h, d, S = f ** 1.5 * ph, f ** 0.5 * ph, np.ones_like(f)
    ll = O.loglike_internal(d, S, f, h, weights=w, gaussleg=True)
    want = -0.5 * (prim(3.) - 2 * prim(2.))
    assert abs(ll / want - 1) < 1e-13
    # log10 grid, the amplitude and noise powers of a real signal (a = -7/6, c = -4.14)
    u, wu = _gl(np.log10(f1), np.log10(f2), 64)
    f = 10 ** u
    ph = np.exp(1j * 3.1 * f)
    h, d, S = f ** (-7 / 6) * ph, 0.7 * f ** (-1.0) * ph, f ** -4.14
    ll = O.loglike_internal(d, S, f, h, weights=wu, log10F=True, gaussleg=True)
    want = -0.5 * (prim(-7 / 3 + 4.14) - 2 * 0.7 * prim(-7 / 6 - 1.0 + 4.14))
    assert abs(ll / want - 1) < 1e-12
    assert abs(O.loglike_internal(d, S, f, h, weights=wu, log10F=False, gaussleg=True) / want - 1) > 0.1
    # the Fisher elements of the same rule: F_ij = 4 int Re(h_i h_j*) / S df
    D = np.vstack([f ** (-7 / 6) * ph, 1j * f ** (-1 / 2) * ph, f ** (-5 / 6) * np.exp(1j * 0.2) * ph])
    F = O.fisher_elements(f, D, S, weights=wu, log10F=True, gaussleg=True)
    pw = [-7 / 6, -1 / 2, -5 / 6]
    ph0 = [1, 1j, np.exp(1j * 0.2)]
    for i in range(3):
        for j in range(3):
            want = (ph0[i] * np.conj(ph0[j])).real * prim(pw[i] + pw[j] + 4.14)
            assert abs(F[i, j] - want) < 1e-12 * np.sqrt(abs(F[i, i] * F[j, j]))
    # Simpson on an equally spaced grid reproduces the engine-independent Fisher elements
    fs = np.linspace(f1, f2, 2001)
    Ds = np.vstack([fs ** (-7 / 6), fs ** (-1 / 2)]).astype(complex)
    Fs = O.fisher_elements(fs, Ds, fs ** -4.14)
    assert abs(Fs[0, 1] / prim(-7 / 6 - 1 / 2 + 4.14) - 1) < 1e-6
