"""Synthetic inputs of the benchmark configurations (numpy only).

PSD: analytic fit S_n(f) = 1e-49 [x^-4.14 - 5 x^-2 + 111 (1 - x^2 + x^4/2)/(1 + x^2/2)],
x = f / 215 Hz (one-sided PSD, passed as `psd`).  Data d = h_inj + n with
Re n, Im n ~ N(0, S_n / (4 df)) drawn from Philox(seed = 20261202 + 1000*config + event).
Detectors: Hanford, Livingston, Virgo.  gmst = GMST(GPS 1126259462.4) in radians
(util.cpp:1797-1846), f_ref = 20 Hz, shift_time = shift_phase = true (the MCMC
convention of mcmc_gw.cpp:2500-2502).
"""
import math

import numpy as np

from . import IMRPHENOMD, PARAMS_DTYPE, params_array

SEED0 = 20261202
MSUN_SEC = 4.925491025543575903411922162094833998e-6


def gps_to_gmst_radian(gps):
    """Greenwich mean sidereal time in radians (the reference's approximation, util.cpp:1797-1846)."""
    J2000 = 2451545.0
    JD = J2000 + (gps - 630763213.) / 86400.
    JD0 = math.floor(JD) + .5
    H = (JD - JD0) * 24
    D0, D = JD0 - J2000, JD - J2000
    tau = D / 36525.
    g = 6.697374558 + 0.06570982441908 * D0 + 1.00273790935 * H + 0.000026 * tau * tau
    hours = float(int(math.floor(g)) % 24)
    return (hours + (g - math.floor(g))) * math.pi / 12.


GMST = gps_to_gmst_radian(1126259462.4)

INJECTION = dict(mass1=36., mass2=29., spin1=(0, 0, 0.3), spin2=(0, 0, -0.1), DL=410., tc=0., phiRef=0.,
                 RA=1.37, DEC=-1.26, psi=2.66, incl=2.9, f_ref=20., gmst=GMST)


def analytic_psd(f):
    x = np.asarray(f, dtype=np.float64) / 215.
    return 1e-49 * (x ** -4.14 - 5 * x ** -2 + 111 * (1 - x ** 2 + x ** 4 / 2) / (1 + x ** 2 / 2))


def uniform_grid(f0, f1, df):
    n = int(round((f1 - f0) / df)) + 1
    return f0 + df * np.arange(n, dtype=np.float64)


def rng(config, event=0, stream=0):
    return np.random.Generator(np.random.Philox(key=SEED0 + 1000 * config + event, counter=stream))


def noise(f, psd, g):
    df = f[1] - f[0]
    sig = np.sqrt(psd / (4 * df))
    return sig * g.standard_normal(len(f)) + 1j * sig * g.standard_normal(len(f))


def fill(arr, **kw):
    for k, v in kw.items():
        arr[k] = v
    return arr


def injection_params(n=1, **over):
    a = params_array(n)
    kw = dict(INJECTION)
    kw.update(over)
    return fill(a, **kw)


def chirp_q_to_m1m2(mc, q):
    """component masses from chirp mass and q = m2/m1 <= 1"""
    eta = q / (1 + q) ** 2
    M = mc / eta ** 0.6
    m1 = M / (1 + q)
    return m1, M - m1


def config0(n_inst=1024, response_fn=None):
    """Plain method, small: one Hanford-like detector, 4017 bins 20-1024 Hz, df = 0.25 Hz."""
    f = uniform_grid(20., 1024., 0.25)
    psd = analytic_psd(f)
    g = rng(0, 0)
    inj = injection_params(1)[0]
    draws = rng(0, 0, stream=1)
    mc = draws.uniform(20, 40, n_inst)
    q = draws.uniform(0.3, 1, n_inst)
    m1, m2 = chirp_q_to_m1m2(mc, q)
    P = injection_params(n_inst)
    P["mass1"], P["mass2"] = m1, m2
    P["spin1"][:, 2] = draws.uniform(-0.9, 0.9, n_inst)
    P["spin2"][:, 2] = draws.uniform(-0.9, 0.9, n_inst)
    P["DL"] = np.exp(draws.uniform(np.log(100), np.log(1000), n_inst))
    return dict(f=f, psd=psd, noise_gen=g, injection=inj, params=P, dets=[0])


def config1(n_temps=16, n_chains=32):
    """Batched PT ensemble: H/L/V, 20-2048 Hz at df = 1/8 Hz (16225 bins), 16 temperatures x 32 chains."""
    f = uniform_grid(20., 2048., 0.125)
    psd = analytic_psd(f)
    temps = 100.0 ** (np.arange(n_temps) / (n_temps - 1))
    return dict(f=f, psd=psd, dets=[0, 1, 2], temps=temps, n_chains=n_chains, injection=injection_params(1)[0])


def make_data(f, psd, dets, inj, response_fn, config, event=0):
    """Injected signal plus seeded Gaussian noise for every detector.  response_fn(f, p, det_index) must return
    the detector response of the injection including its time-of-flight phase."""
    g = rng(config, event)
    data = []
    for d in range(len(dets)):
        h = response_fn(f, inj, d)
        data.append(h + noise(f, psd, g))
    return data


def perturbed_params(base, n, step, config=1, scale=1.0):
    """n parameter vectors scattered around `base` (a PARAMS_DTYPE record), seeded by step."""
    g = rng(config, 0, stream=1000 + step)
    P = params_array(n)
    for name in PARAMS_DTYPE.names:
        P[name] = base[name]
    mc0 = (base["mass1"] * base["mass2"]) ** 0.6 / (base["mass1"] + base["mass2"]) ** 0.2
    q0 = base["mass2"] / base["mass1"]
    mc = mc0 * np.exp(scale * 0.05 * g.standard_normal(n))
    q = np.clip(q0 + scale * 0.1 * g.standard_normal(n), 0.2, 1.0)
    m1, m2 = chirp_q_to_m1m2(mc, q)
    P["mass1"], P["mass2"] = m1, m2
    P["spin1"][:, 2] = np.clip(base["spin1"][2] + scale * 0.1 * g.standard_normal(n), -0.95, 0.95)
    P["spin2"][:, 2] = np.clip(base["spin2"][2] + scale * 0.1 * g.standard_normal(n), -0.95, 0.95)
    P["DL"] = base["DL"] * np.exp(scale * 0.1 * g.standard_normal(n))
    P["RA"] = base["RA"] + scale * 0.1 * g.standard_normal(n)
    P["DEC"] = np.clip(base["DEC"] + scale * 0.1 * g.standard_normal(n), -1.5, 1.5)
    P["psi"] = base["psi"] + scale * 0.1 * g.standard_normal(n)
    P["incl"] = np.clip(base["incl"] + scale * 0.1 * g.standard_normal(n), 0.01, 3.13)
    P["phiRef"] = base["phiRef"] + scale * 0.2 * g.standard_normal(n)
    P["tc"] = base["tc"] + scale * 0.002 * g.standard_normal(n)
    return P


def precessing_params(n, config=3, seed_stream=0):
    """IMRPhenomPv2 parameter vectors: |chi| ~ U(0, 0.9) with isotropic directions around the injection."""
    g = rng(config, 0, stream=seed_stream)
    P = perturbed_params(injection_params(1)[0], n, seed_stream, config=config, scale=0.5)
    for s in ("spin1", "spin2"):
        mag = g.uniform(0, 0.9, n)
        ct = g.uniform(-1, 1, n)
        ph = g.uniform(0, 2 * np.pi, n)
        st = np.sqrt(1 - ct * ct)
        P[s][:, 0] = mag * st * np.cos(ph)
        P[s][:, 1] = mag * st * np.sin(ph)
        P[s][:, 2] = mag * ct
    P["chip"] = -1.0
    return P


def config2(n_events=256, chains=8):
    """Uneven batch: n_events sources with total mass log-uniform 5-200 Msun, q = m2/m1 uniform 0.1-1,
    aligned spins uniform +-0.95, durations log-uniform 4-128 s; each grid runs from 20 Hz to
    min(2048, 0.3 / (M G/c^3)) Hz at df = 1/T, with the analytic PSD on H/L/V.  Extrinsic parameters,
    distance and reference frequency are those of the fixed injection.  Returns the event specs and the
    `chains` parameter vectors per event (scattered around each injection)."""
    g = rng(2, 0, stream=1)
    M = np.exp(g.uniform(np.log(5.), np.log(200.), n_events))
    q = g.uniform(0.1, 1.0, n_events)
    c1 = g.uniform(-0.95, 0.95, n_events)
    c2 = g.uniform(-0.95, 0.95, n_events)
    T = np.exp(g.uniform(np.log(4.), np.log(128.), n_events))
    events = []
    for e in range(n_events):
        df = 1.0 / T[e]
        fmax = min(2048., 0.3 / (M[e] * MSUN_SEC))
        f = 20. + df * np.arange(int(np.floor((fmax - 20.) / df)) + 1)
        inj = injection_params(1, mass1=M[e] / (1 + q[e]), mass2=M[e] * q[e] / (1 + q[e]),
                               spin1=(0., 0., c1[e]), spin2=(0., 0., c2[e]))[0]
        P = perturbed_params(inj, chains, step=e, config=2, scale=0.05)
        events.append(dict(f=f, psd=analytic_psd(f), dets=[0, 1, 2], injection=inj, params=P, T=T[e], M=M[e]))
    return events
