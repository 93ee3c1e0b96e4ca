"""Parallel-tempered ensemble driver over the batched engine (the caller side of the hot path).

It mirrors what the reference's sampler does with the likelihood of every chain in one step:
Gaussian proposals in the MCMC_IMRPhenomD basis (RA, sin DEC, psi, cos incl, phiRef, tc, ln DL,
ln Mc, eta, chi1, chi2), a uniform prior box, the Metropolis-Hastings rule of
mcmc_sampler_internals.cpp:96-142 (reject iff (ll' - ll) / T < log U; a NaN likelihood rejects) and,
every `swap_every` steps, the adjacent-temperature swap rule of mcmc_sampler_internals.cpp:1121-1190
(accept iff exp((ll1 - ll2) / T2 - (ll1 - ll2) / T1) >= U).  All random numbers of a step come from
one Philox stream keyed by (seed, step), so a run in FP64 and a run in mixed precision consume
identical streams.

Guard band.  In mixed precision every accept/reject and swap decision whose margin lies within
2 * tol / T (tol = the mixed-precision likelihood bound, 1e-3) of the threshold is re-decided with
FP64 likelihoods of both states, so the decision sequence equals the all-FP64 run's whenever the
mixed error stays within tol.  The number of such re-evaluations is reported.

Likelihoods are evaluated by the device engine on one batch per step (all temperatures x chains);
parameter vectors cross to the device and log-likelihoods back every step.
"""
import ctypes
import time

import numpy as np

from . import FP64, MIXED, IMRPHENOMD, PARAMS_DTYPE, check, lib
from .engine import DeviceArray, loglike_batch

DIM = 11
NAMES = ("RA", "sinDEC", "psi", "cos_incl", "phiRef", "tc", "lnDL", "lnMc", "eta", "chi1", "chi2")
# uniform prior box in the MCMC basis
LOWER = np.array([0., -1., 0., -1., 0., -0.1, np.log(10.), np.log(5.), 0.05, -0.95, -0.95])
UPPER = np.array([2 * np.pi, 1., np.pi, 1., 2 * np.pi, 0.1, np.log(5000.), np.log(100.), 0.25, 0.95, 0.95])


def theta_from_params(rec):
    """the MCMC_IMRPhenomD basis vector of a PARAMS_DTYPE record (fisher.cpp:1841-2161 layout)"""
    m1, m2 = float(rec["mass1"]), float(rec["mass2"])
    mc = (m1 * m2) ** 0.6 / (m1 + m2) ** 0.2
    eta = m1 * m2 / (m1 + m2) ** 2
    return np.array([rec["RA"], np.sin(rec["DEC"]), rec["psi"], np.cos(rec["incl"]), rec["phiRef"], rec["tc"],
                     np.log(rec["DL"]), np.log(mc), eta, rec["spin1"][2], rec["spin2"][2]], dtype=np.float64)


def params_from_theta(theta, template):
    """vectorised repack of MCMC_IMRPhenomD basis vectors (n x 11) into PARAMS_DTYPE records; the masses
    follow calculate_mass1/2 (util.cpp:1520-1540)"""
    theta = np.atleast_2d(theta)
    n = len(theta)
    P = np.zeros(n, dtype=PARAMS_DTYPE)
    for k in PARAMS_DTYPE.names:
        P[k] = template[k]
    mc = np.exp(theta[:, 7])
    eta = np.clip(theta[:, 8], 1e-6, 0.25)
    e35 = eta ** 0.6
    root = np.sqrt(np.maximum(1. - 4. * eta, 0.))
    P["mass1"] = 0.5 * (mc / e35 + root * mc / e35)
    P["mass2"] = 0.5 * (mc / e35 - root * mc / e35)
    P["RA"] = theta[:, 0]
    P["DEC"] = np.arcsin(np.clip(theta[:, 1], -1, 1))
    P["psi"] = theta[:, 2]
    P["incl"] = np.arccos(np.clip(theta[:, 3], -1, 1))
    P["phiRef"] = theta[:, 4]
    P["tc"] = theta[:, 5]
    P["DL"] = np.exp(theta[:, 6])
    P["spin1"][:, 2] = theta[:, 9]
    P["spin2"][:, 2] = theta[:, 10]
    return P


def in_prior(theta):
    return np.all((theta >= LOWER) & (theta <= UPPER), axis=1)


def step_stream(seed, step):
    return np.random.Generator(np.random.Philox(key=np.array([seed, step], dtype=np.uint64)))


class PTEnsemble:
    """n_temps x n_chains chains on one resident event; instance index = temp * n_chains + chain."""

    def __init__(self, event, template, theta0, sigma, temps, n_chains, precision=FP64, model=IMRPHENOMD,
                 seed=20261203, swap_every=10, tol=1e-3, guard=True):
        self.event, self.template, self.model = event, template, model
        self.temps = np.asarray(temps, dtype=np.float64)
        self.nt, self.nc = len(self.temps), int(n_chains)
        self.n = self.nt * self.nc
        self.T = np.repeat(self.temps, self.nc)
        self.sigma = np.asarray(sigma, dtype=np.float64)
        self.precision, self.seed, self.swap_every, self.tol = precision, int(seed), int(swap_every), tol
        self.guard = guard and precision == MIXED
        self.theta = np.tile(np.asarray(theta0, dtype=np.float64), (self.n, 1))
        self.d_params = self.d_logl = None
        self.ll = self._eval(self.theta)
        self.guard_evals = 0
        self.history = {"accept": [], "swap": [], "proposed_ll": []}

    # -- likelihood -----------------------------------------------------------
    def _eval(self, theta, precision=None):
        P = params_from_theta(theta, self.template)
        prec = self.precision if precision is None else precision
        if len(P) == self.n:
            if self.d_params is None:
                self.d_params = DeviceArray(self.n * PARAMS_DTYPE.itemsize)
                self.d_logl = DeviceArray(self.n * 8)
            out = np.zeros(self.n)
            check(lib().gwh_memcpy_h2d(self.d_params.ptr, P.ctypes.data_as(ctypes.c_void_p), P.nbytes))
            check(lib().gwh_loglike_batch_device(None, self.event.id, self.model, prec, self.d_params.ptr, self.n,
                                                 self.d_logl.ptr, None))
            check(lib().gwh_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.d_logl.ptr, out.nbytes))
            return out
        return loglike_batch(self.event, P, self.model, prec)

    def _fp64(self, thetas):
        self.guard_evals += len(thetas)
        return loglike_batch(self.event, params_from_theta(np.asarray(thetas), self.template), self.model, FP64)

    # -- one step -------------------------------------------------------------
    def step(self, s):
        g = step_stream(self.seed, s)
        z = g.standard_normal((self.n, DIM))
        logu = np.log(g.random(self.n))
        prop = self.theta + z * self.sigma[None, :] * np.sqrt(self.T)[:, None]
        ok = in_prior(prop)
        llp = self._eval(np.where(ok[:, None], prop, self.theta))
        llp = np.where(ok, llp, np.nan)
        self.history["proposed_ll"].append(np.where(ok, llp, np.nan))
        ratio = np.where(np.isnan(llp), -np.inf, (llp - self.ll) / self.T)
        margin = ratio - logu
        accept = margin >= 0
        if self.guard:
            band = np.isfinite(margin) & (np.abs(margin) < 2 * self.tol / self.T)
            idx = np.nonzero(band)[0]
            if len(idx):
                v = self._fp64(np.concatenate([prop[idx], self.theta[idx]]))
                r64 = (v[:len(idx)] - v[len(idx):]) / self.T[idx]
                accept[idx] = (r64 - logu[idx]) >= 0
        self.theta = np.where(accept[:, None], prop, self.theta)
        self.ll = np.where(accept, llp, self.ll)
        self.history["accept"].append(accept)
        swaps = np.zeros((self.nt - 1, self.nc), dtype=bool)
        if self.swap_every and (s + 1) % self.swap_every == 0:
            logalpha = np.log(g.random((self.nt - 1, self.nc)))
            for k in range(self.nt - 1):
                i1 = k * self.nc + np.arange(self.nc)
                i2 = i1 + self.nc
                T1, T2 = self.temps[k], self.temps[k + 1]
                d = self.ll[i1] - self.ll[i2]
                m = d / T2 - d / T1 - logalpha[k]
                acc = m >= 0
                if self.guard:
                    band = np.abs(m) < 2 * self.tol * abs(1 / T2 - 1 / T1)
                    idx = np.nonzero(band)[0]
                    if len(idx):
                        v = self._fp64(np.concatenate([self.theta[i1[idx]], self.theta[i2[idx]]]))
                        d64 = v[:len(idx)] - v[len(idx):]
                        acc[idx] = (d64 / T2 - d64 / T1 - logalpha[k][idx]) >= 0
                a1, a2 = i1[acc], i2[acc]
                self.theta[a1], self.theta[a2] = self.theta[a2].copy(), self.theta[a1].copy()
                self.ll[a1], self.ll[a2] = self.ll[a2].copy(), self.ll[a1].copy()
                swaps[k] = acc
        self.history["swap"].append(swaps)

    def run(self, steps):
        t0 = time.perf_counter()
        for s in range(steps):
            self.step(s)
        return time.perf_counter() - t0


def proposal_scale(event, template, theta0, fisher_fn, factor=0.5):
    """per-parameter proposal widths factor / sqrt(F_ii) from the network Fisher matrix at theta0"""
    F = fisher_fn(event, params_from_theta(theta0, template))
    d = np.diag(F)
    s = factor / np.sqrt(np.where(d > 0, d, np.inf))
    return np.where(np.isfinite(s) & (s > 0), s, 1e-3)
