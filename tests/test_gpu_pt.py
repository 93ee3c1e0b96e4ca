"""The parallel-tempered ensemble on the GPU (16 temperatures x 32 chains on the 3-detector, 16225-bin
grid): the mixed-precision run with its FP64 guard band makes exactly the FP64 run's accept and swap
decisions, and every proposed log-likelihood agrees with FP64 within 1e-3."""
import numpy as np
import pytest

import gw_analysis_tools_amd as G
from gw_analysis_tools_amd import pt
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu


def _setup():
    cfg = W.config1()
    f, psd, dets = cfg["f"], cfg["psd"], cfg["dets"]
    inj = W.injection_params(1)[0]
    pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item())
                            for k in inj.dtype.names})
    resp = G.create_coherent_GW_detection(f, pinj, dets)
    g = W.rng(1, 0)
    data = [r + W.noise(f, psd, g) for r in resp]
    ev = G.Event(f, data, [psd] * 3, dets)
    th0 = pt.theta_from_params(inj)
    sig = pt.proposal_scale(ev, inj, th0, lambda e, P: G.fisher_batch(e, P, G.IMRPHENOMD, G.BASIS_PHENOMD_MCMC,
                                                                        11)[0])
    return ev, inj, th0, sig, cfg["temps"]


def test_pt_mixed_decisions_equal_fp64():
    ev, inj, th0, sig, temps = _setup()
    steps = 60
    a = pt.PTEnsemble(ev, inj, th0, sig, temps, 32, precision=G.FP64)
    a.run(steps)
    b = pt.PTEnsemble(ev, inj, th0, sig, temps, 32, precision=G.MIXED)
    b.run(steps)
    acc = np.array(a.history["accept"])
    print("acceptance T=1 %.3f, all %.3f; swaps %d; guard re-evaluations %d" %
          (acc[:, :32].mean(), acc.mean(), int(np.sum(a.history["swap"])), b.guard_evals))
    assert 0.02 < acc.mean() < 0.98
    for x, y in zip(a.history["accept"], b.history["accept"]):
        assert np.array_equal(x, y)
    for x, y in zip(a.history["swap"], b.history["swap"]):
        assert np.array_equal(x, y)
    la, lb = np.array(a.history["proposed_ll"]), np.array(b.history["proposed_ll"])
    fin = np.isfinite(la)
    assert np.array_equal(fin, np.isfinite(lb))
    d = np.abs(la[fin] - lb[fin])
    print("proposed logL mixed vs FP64: max %.3e median %.3e over %d" % (d.max(), np.median(d), d.size))
    assert d.max() <= 1e-3
    ev.close()
