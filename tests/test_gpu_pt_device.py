"""Parallel tempering with the whole step on the device, replayed from one HIP graph per step
(gwh_pt_*), and the captured likelihood step (gwh_loglike_graph_*).

* FP64 device run vs the CPU restatement (oracle/pt_oracle.py) driven by the C oracle's likelihood:
  the same accept and swap decisions every step, states equal to the last bits of the normal deviates,
  log-likelihoods within the FP64 bar (1e-9 relative).
* mixed precision with the guard band vs the FP64 device run: identical decisions and bitwise identical
  states (the proposals do not depend on the likelihood precision).
* replays of a captured likelihood step equal eager launches bit for bit, also after the parameters change.
"""
import ctypes

import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from oracle import pt_oracle as PO
from gw_analysis_tools_amd import pt
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu

LL_TOL = 1e-9
# step widths in the MCMC_IMRPhenomD basis (RA, sinDEC, psi, cos incl, phiRef, tc, lnDL, lnMc, eta, chi1, chi2)
WIDTHS = np.array([0.02, 0.02, 0.05, 0.03, 0.05, 2e-4, 0.03, 2e-4, 2e-3, 0.05, 0.05])


def _small_event(df=0.25, fmax=512.):
    f = W.uniform_grid(20., fmax, df)
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    inj = W.injection_params(1)[0]
    pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item())
                            for k in inj.dtype.names})
    resp = G.create_coherent_GW_detection(f, pinj, dets)
    g = W.rng(1, 0)
    data = [r + W.noise(f, psd, g) for r in resp]
    return f, psd, dets, data, inj


def _oracle_ll(f, data, psd, dets, inj):
    tmpl = O.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item())
                            for k in inj.dtype.names})

    def ll(thetas):
        plist = [O.repack_parameters(t, "MCMC_IMRPhenomD", 11, tmpl) for t in np.atleast_2d(thetas)]
        out, _ = O.network_loglike_batch(f, data, [psd] * 3, dets, plist, O.IMRPHENOMD, nthreads=8)
        return out
    return ll


def test_device_pt_fp64_matches_oracle_restatement():
    f, psd, dets, data, inj = _small_event()
    ev = G.Event(f, data, [psd] * 3, dets)
    th0 = pt.theta_from_params(inj)
    temps = np.array([1.0, 1.6, 2.6, 4.0])
    nc, steps, seed, swap_every = 8, 24, 424242, 3
    dev = G.DevicePT(ev, inj, th0, WIDTHS, temps, nc, pt.LOWER, pt.UPPER, precision=G.FP64, seed=seed,
                     swap_every=swap_every)
    ref = PO.PTOracle(_oracle_ll(f, data, psd, dets, inj), th0, WIDTHS, temps, nc, pt.LOWER, pt.UPPER, seed,
                      swap_every)
    th, ll, st = dev.state()
    assert st == 0 and np.all(np.abs(ll - ref.ll) <= LL_TOL * np.abs(ref.ll))
    n_acc = n_swap = 0
    for s in range(steps):
        dev.run(1)
        ref.step()
        prop, llp, acc, sw = dev.last_step()
        h = ref.history[-1]
        assert np.array_equal(acc, h["accept"]), s
        assert np.array_equal(sw, h["swapped"]), s
        assert np.allclose(prop, h["prop"], rtol=1e-13, atol=1e-15), s
        fin = np.isfinite(h["llp"])
        assert np.array_equal(fin, np.isfinite(llp)), s
        assert np.all(np.abs(llp[fin] - h["llp"][fin]) <= LL_TOL * np.abs(h["llp"][fin])), s
        n_acc += acc.sum()
        n_swap += sw.sum()
    th, ll, st = dev.state()
    assert st == steps
    assert np.allclose(th, ref.theta, rtol=1e-13, atol=1e-15)
    assert np.all(np.abs(ll - ref.ll) <= LL_TOL * np.abs(ref.ll))
    acc_count, swap_count, guard = dev.counters()
    assert acc_count.sum() == n_acc and swap_count.sum() == n_swap and guard == 0
    print("device PT vs oracle: %d steps, %d accepts, %d swaps" % (steps, n_acc, n_swap))
    assert 0 < n_acc < steps * len(temps) * nc and n_swap > 0
    dev.close()
    ev.close()


def _c1_event():
    cfg = W.config1()
    f, psd, dets = cfg["f"], cfg["psd"], cfg["dets"]
    inj = W.injection_params(1)[0]
    pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item())
                            for k in inj.dtype.names})
    resp = G.create_coherent_GW_detection(f, pinj, dets)
    g = W.rng(1, 0)
    data = [r + W.noise(f, psd, g) for r in resp]
    return G.Event(f, data, [psd] * 3, dets), inj, cfg["temps"]


@pytest.mark.parametrize("tol", [1e-3, 0.05])
def test_device_pt_mixed_guard_takes_fp64_decisions(tol):
    """tol 0.05 widens the band far beyond the mixed error, so many decisions go through the FP64
    re-evaluation inside the graph and must still equal the FP64 run's"""
    ev, inj, temps = _c1_event()
    th0 = pt.theta_from_params(inj)
    nc, steps = 16, 40
    runs = {}
    for prec in (G.FP64, G.MIXED):
        d = G.DevicePT(ev, inj, th0, WIDTHS, temps, nc, pt.LOWER, pt.UPPER, precision=prec, seed=99,
                       swap_every=4, tol=tol)
        hist = []
        for s in range(steps):
            d.run(1)
            prop, llp, acc, sw = d.last_step()
            hist.append((acc, sw, llp))
        runs[prec] = (hist, d.state(), d.counters())
        d.close()
    (ha, sa, ca), (hb, sb, cb) = runs[G.FP64], runs[G.MIXED]
    for s, (x, y) in enumerate(zip(ha, hb)):
        assert np.array_equal(x[0], y[0]), s
        assert np.array_equal(x[1], y[1]), s
        fin = np.isfinite(x[2])
        assert np.array_equal(fin, np.isfinite(y[2]))
        assert np.all(np.abs(x[2][fin] - y[2][fin]) <= 1e-3)
    assert np.array_equal(sa[0], sb[0])  # states bit for bit
    assert np.all(np.abs(sa[1] - sb[1]) <= 1e-3)
    assert np.array_equal(ca[0], cb[0]) and np.array_equal(ca[1], cb[1])
    print("tol %g: accepts %d, swaps %d, guard FP64 evaluations %d" % (tol, ca[0].sum(), ca[1].sum(), cb[2]))
    assert ca[0].sum() > 0 and ca[1].sum() > 0
    if tol > 1e-2:
        assert cb[2] > 0
    ev.close()


def test_device_pt_graph_of_several_steps_and_reproducibility():
    ev, inj, temps = _c1_event()
    th0 = pt.theta_from_params(inj)
    kw = dict(precision=G.MIXED, seed=5, swap_every=2)
    a = G.DevicePT(ev, inj, th0, WIDTHS, temps[:4], 8, pt.LOWER, pt.UPPER, steps_per_graph=5, **kw)
    b = G.DevicePT(ev, inj, th0, WIDTHS, temps[:4], 8, pt.LOWER, pt.UPPER, steps_per_graph=1, **kw)
    a.run(20)
    for _ in range(20):
        b.run(1, sync=False)
    sa, sb = a.state(), b.state()
    assert sa[2] == sb[2] == 20
    assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])
    with pytest.raises(G.GWHError):
        a.run(3)  # not a multiple of steps_per_graph
    # the event is held while the runs exist
    assert G.lib().gwh_event_destroy(ev.id) == G.GWH_ERR_BUSY
    a.close()
    b.close()
    ev.close()


def test_device_pt_argument_errors():
    ev, inj, temps = _c1_event()
    th0 = pt.theta_from_params(inj)
    with pytest.raises(G.GWHError):
        G.DevicePT(ev, inj, th0, WIDTHS, np.ones(65), 2, pt.LOWER, pt.UPPER)  # too many temperatures
    with pytest.raises(G.GWHError):
        G.DevicePT(ev, inj, th0, WIDTHS, [1.0, -1.0], 2, pt.LOWER, pt.UPPER)  # bad temperature
    with pytest.raises(G.GWHError):
        G.DevicePT(ev, inj, th0, WIDTHS, [1.0], 2, pt.LOWER, pt.UPPER, basis=G.BASIS_PHENOMD_FULL)
    with pytest.raises(G.GWHError):
        G.DevicePT(12345, inj, th0, WIDTHS, [1.0], 2, pt.LOWER, pt.UPPER)  # unknown event
    ev.close()


def test_loglike_graph_replays_equal_eager_launches():
    ev, inj, temps = _c1_event()
    n = 96
    base = W.perturbed_params(inj, n, step=3)
    d_par = G.DeviceArray(n * G.PARAMS_DTYPE.itemsize)
    d_ll = G.DeviceArray(n * 8)
    for prec in (G.FP64, G.MIXED):
        g = G.LoglikeGraph(ev, d_par.ptr, n, d_ll.ptr, precision=prec)
        for rep in range(3):
            P = base.copy()
            P["DL"] *= 1.0 + 0.05 * rep
            P["tc"] += 1e-4 * rep
            ref = G.loglike_batch(ev, P, G.IMRPHENOMD, prec)
            G.check(G.lib().gwh_memcpy_h2d(d_par.ptr, P.ctypes.data_as(ctypes.c_void_p), P.nbytes))
            g.launch()
            g.synchronize()
            out = d_ll.to_host(np.float64, n)
            assert np.array_equal(out, ref), (prec, rep)
        assert G.lib().gwh_event_destroy(ev.id) == G.GWH_ERR_BUSY
        g.close()
    d_par.free()
    d_ll.free()
    ev.close()


def test_loglike_graph_multi_event():
    specs = W.config2(n_events=6, chains=4)
    g = W.rng(2, 1)
    evs, P, ids = [], [], []
    for sp in specs:
        f, psd, inj = sp["f"], sp["psd"], sp["injection"]
        pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item())
                                for k in inj.dtype.names})
        data = [r + W.noise(f, psd, g) for r in G.create_coherent_GW_detection(f, pinj, sp["dets"])]
        e = G.Event(f, data, [psd] * 3, sp["dets"])
        evs.append(e)
        P.append(sp["params"])
        ids += [e.id] * len(sp["params"])
    P, ids = np.concatenate(P), np.array(ids, dtype=np.int32)
    n = len(P)
    ref = G.loglike_batch(int(ids[0]), P, G.IMRPHENOMD, G.FP64, events=list(ids))
    d_par = G.DeviceArray.from_host(P)
    d_ev = G.DeviceArray.from_host(ids)
    d_ll = G.DeviceArray(n * 8)
    gr = G.LoglikeGraph(None, d_par.ptr, n, d_ll.ptr, precision=G.FP64, events=list(ids), d_events=d_ev.ptr)
    for _ in range(2):
        gr.launch()
        gr.synchronize()
        assert np.array_equal(d_ll.to_host(np.float64, n), ref)
    gr.close()
    for a in (d_par, d_ev, d_ll):
        a.free()
    for e in evs:
        e.close()


def test_device_pt_full_c1_shape_identical_decisions():
    """The benchmark's tempering run at its full shape: 16 temperatures x 32 chains on the 16225-bin H/L/V event,
    2000 steps, ladder swaps every 10th step, 10 steps per graph replay.  The mixed engine with its FP64 guard band
    must take exactly the FP64 run's decisions: equal acceptance and swap counts per chain and a bit-identical
    final state (parameters), with logL within the mixed tolerance."""
    ev, inj, temps = _c1_event()
    th0 = pt.theta_from_params(inj)
    runs = {}
    for prec in (G.FP64, G.MIXED):
        d = G.DevicePT(ev, inj, th0, WIDTHS, temps, 32, pt.LOWER, pt.UPPER, precision=prec, seed=2026,
                       swap_every=10, steps_per_graph=10)
        d.run(2000)
        runs[prec] = (d.state(), d.counters())
        d.close()
    (sa, ca), (sb, cb) = runs[G.FP64], runs[G.MIXED]
    print("2000 steps x 512 chains: accepts %d, swaps %d, guard FP64 evaluations %d"
          % (ca[0].sum(), ca[1].sum(), cb[2]))
    assert np.array_equal(ca[0], cb[0]) and np.array_equal(ca[1], cb[1])
    assert np.array_equal(sa[0], sb[0])
    assert np.all(np.abs(sa[1] - sb[1]) <= 1e-3)
    assert ca[0].sum() > 1000 and ca[1].sum() > 0
    ev.close()
