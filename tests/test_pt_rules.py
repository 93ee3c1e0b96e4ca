"""The parallel-tempering driver's decision logic on the CPU with a stand-in likelihood: MH and swap
rules as in the reference sampler, identical random streams per step, and the mixed-precision guard
band reproducing the FP64 decision sequence exactly when the likelihood error stays within tol."""
import numpy as np

from gw_analysis_tools_amd import pt, FP64, MIXED
from gw_analysis_tools_amd import workloads as W


class FakeEngine(pt.PTEnsemble):
    """Gaussian log-likelihood in the MCMC basis; 'mixed' adds a bounded, state-dependent error."""
    centre = None
    err = 0.0

    def _exact(self, theta):
        d = (np.atleast_2d(theta) - self.centre) / (10 * self.sigma)
        return 50.0 - 0.5 * np.sum(d * d, axis=1)

    def _eval(self, theta, precision=None):
        v = self._exact(theta)
        prec = self.precision if precision is None else precision
        if prec == MIXED:
            # deterministic pseudo-noise in [-err, err]
            v = v + self.err * np.sin(1e3 * np.sum(np.atleast_2d(theta), axis=1))
        return v

    def _fp64(self, thetas):
        self.guard_evals += len(thetas)
        return self._exact(np.asarray(thetas))


def _run(precision, err, guard=True, steps=120):
    tmpl = W.injection_params(1)[0]
    th0 = pt.theta_from_params(tmpl)
    FakeEngine.centre = th0
    FakeEngine.err = err
    sig = np.full(pt.DIM, 1e-3)
    e = FakeEngine(None, tmpl, th0, sig, temps=100.0 ** (np.arange(4) / 3), n_chains=8, precision=precision,
                   seed=7, swap_every=5, tol=1e-3, guard=guard)
    e.run(steps)
    return e


def test_mh_and_swap_rules():
    e = _run(FP64, 0.0, steps=60)
    acc = np.array(e.history["accept"])
    assert 0.05 < acc.mean() < 0.95
    sw = np.array(e.history["swap"])
    assert sw.shape == (60, 3, 8) and sw[4].any() and not sw[0].any()  # swaps only every 5th step
    # the stored likelihoods are those of the stored states
    assert np.allclose(e.ll, e._exact(e.theta))


def test_guard_band_reproduces_fp64_decisions():
    a = _run(FP64, 0.0)
    b = _run(MIXED, 9e-4)  # error inside the guaranteed bound
    assert all(np.array_equal(x, y) for x, y in zip(a.history["accept"], b.history["accept"]))
    assert all(np.array_equal(x, y) for x, y in zip(a.history["swap"], b.history["swap"]))
    assert np.allclose(a.theta, b.theta, rtol=0, atol=0)
    assert b.guard_evals > 0


def test_without_guard_decisions_can_diverge():
    a = _run(FP64, 0.0)
    c = _run(MIXED, 0.5, guard=False)  # a large error and no guard: some decision flips
    same = all(np.array_equal(x, y) for x, y in zip(a.history["accept"], c.history["accept"]))
    assert not same


def test_repack_round_trip():
    tmpl = W.injection_params(1)[0]
    th = pt.theta_from_params(tmpl)
    P = pt.params_from_theta(th, tmpl)
    assert np.allclose(pt.theta_from_params(P[0]), th, rtol=1e-13, atol=1e-15)
    assert abs(P["mass1"][0] - 36) < 1e-10 and abs(P["mass2"][0] - 29) < 1e-10
