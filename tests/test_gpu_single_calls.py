"""The reference-signature single-call entry points under concurrency, and their event cache modes.

Samplers call MCMC_likelihood_extrinsic from many host threads at once (mcmc_sampler.cpp:4585-4600).  The
engine coalesces such calls into shared launches; every result must be bit-identical to the same call made
alone, whatever the cache mode (content hash, pointer key, registered arrays), and equal to the batched API on
an event built from the same arrays."""
import ctypes
import os
import subprocess
import threading

import numpy as np
import pytest

import gw_analysis_tools_amd as G
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(G.LIB_PATH), "gwh_threaded_calls")
DP = ctypes.POINTER(ctypes.c_double)


class Arrays:
    """the sampler's host arrays (c1 grid, H/L/V, injection + noise) and the pointer tables of one call"""

    def __init__(self, n_params=192):
        c = W.config1()
        self.f, self.psd, self.dets = c["f"], c["psd"], c["dets"]
        inj = c["injection"]
        p = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item()) for k in inj.dtype.names})
        g = W.rng(1, 0)
        self.data = [np.ascontiguousarray(r + W.noise(self.f, self.psd, g), dtype=np.complex128).view(np.float64)
                     for r in G.create_coherent_GW_detection(self.f, p, self.dets)]
        self.psds = [np.ascontiguousarray(self.psd) for _ in self.dets]
        D, n = len(self.dets), len(self.f)
        self.fl = (DP * D)(*[self.f.ctypes.data_as(DP)] * D)
        self.dl = (DP * D)(*[x.ctypes.data_as(DP) for x in self.data])
        self.sl = (DP * D)(*[x.ctypes.data_as(DP) for x in self.psds])
        self.lens = (ctypes.c_int * D)(*[n] * D)
        self.di = (ctypes.c_int * D)(*self.dets)
        self.P = W.perturbed_params(inj, n_params, step=5, config=1, scale=0.3)
        self.recs = [G.Params.from_buffer_copy(self.P[i].tobytes()) for i in range(n_params)]

    def call(self, i):
        return G.lib().gwh_mcmc_likelihood_extrinsic(1, ctypes.byref(self.recs[i]), G.IMRPHENOMD, self.lens, self.fl,
                                                     self.dl, self.sl, None, 0, 0, self.di, len(self.dets), 0.0)

    def register(self):
        h = ctypes.c_int(-1)
        D = len(self.dets)
        G.check(G.lib().gwh_host_data_register(D, self.lens, self.fl, self.dl, self.sl, None, ctypes.byref(h)))
        return h.value


def _threaded(A, threads):
    out = np.full(len(A.recs), np.nan)
    start = threading.Barrier(threads)

    def work(t):
        start.wait()
        for i in range(t, len(A.recs), threads):
            out[i] = A.call(i)

    pool = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for th in pool:
        th.start()
    for th in pool:
        th.join()
    return out


def _stats():
    st = (ctypes.c_double * 9)()
    G.check(G.lib().gwh_single_call_stats(st))
    return list(st)


def test_threaded_calls_equal_sequential_and_the_batched_api():
    A = Arrays()
    G.check(G.lib().gwh_set_single_call_cache(G.CACHE_BY_CONTENT))
    seq = np.array([A.call(i) for i in range(len(A.recs))])
    assert np.all(np.isfinite(seq))
    before = _stats()
    thr = _threaded(A, 32)
    after = _stats()
    print("coalescing: %d shared launches, %d instances, at most %d calls in one launch"
          % (after[0] - before[0], after[1] - before[1], after[3]))
    assert np.array_equal(thr, seq)
    # the batched API on an event built from the same arrays: MCMC_likelihood_extrinsic evaluates t_c = T_eff - tc
    ev = G.Event(A.f, [x.view(np.complex128) for x in A.data], A.psds, A.dets)
    Q = A.P.copy()
    Q["tc"] = 0.0 - A.P["tc"]
    assert np.array_equal(G.loglike_batch(ev, Q, G.IMRPHENOMD, G.FP64), seq)
    ev.close()


def test_pointer_mode_and_registered_arrays_give_the_content_mode_results():
    A = Arrays(n_params=48)
    G.check(G.lib().gwh_set_single_call_cache(G.CACHE_BY_CONTENT))
    ref = np.array([A.call(i) for i in range(len(A.recs))])
    G.check(G.lib().gwh_set_single_call_cache(G.CACHE_BY_POINTER))
    try:
        assert np.array_equal(np.array([A.call(i) for i in range(len(A.recs))]), ref)
        assert np.array_equal(_threaded(A, 16), ref)
    finally:
        G.check(G.lib().gwh_set_single_call_cache(G.CACHE_BY_CONTENT))
    h = A.register()
    try:
        assert np.array_equal(np.array([A.call(i) for i in range(len(A.recs))]), ref)
        assert np.array_equal(_threaded(A, 16), ref)
    finally:
        G.check(G.lib().gwh_host_data_release(h))
    # released: content hashing again, which sees a changed array
    A.data[0][:] *= 0.5
    changed = np.array([A.call(i) for i in range(4)])
    assert not np.array_equal(changed, ref[:4])
    with pytest.raises(G.GWHError):
        G.check(G.lib().gwh_host_data_release(h))
    with pytest.raises(G.GWHError):
        G.check(G.lib().gwh_set_single_call_cache(7))


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="threaded-call driver not built")
@pytest.mark.parametrize("mode", ["content", "pointer", "register"])
def test_cpp_driver_threads_equal_one_thread(tmp_path, mode):
    """the C++ layer called from 32 threads of one process gives the one-thread results bit for bit, with calls
    merged into shared launches"""
    A = Arrays(n_params=256)
    inp = tmp_path / "in.bin"
    G.write_threaded_calls_input(str(inp), A.f, [x.view(np.complex128) for x in A.data], A.psds, A.dets, A.P)
    res = {}
    for t in (1, 32):
        out = tmp_path / ("out%d.bin" % t)
        r = subprocess.run([DRIVER, str(inp), str(out), str(t), mode, "IMRPhenomD"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        print(r.stdout.strip())
        res[t] = (np.fromfile(str(out), dtype=np.float64), r.stdout)
    assert np.all(np.isfinite(res[1][0]))
    assert np.array_equal(res[1][0], res[32][0])
    import json
    st = json.loads(res[32][1])
    assert st["max_calls_per_batch"] > 1
