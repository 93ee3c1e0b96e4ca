"""The allocation-free likelihood entry point (gwh_loglike_batch_device_noalloc) against
gwh_loglike_batch_device through the C ABI: bit-identical results on two calls in a row, the second
smaller, for FP64 and mixed, IMRPhenomD and IMRPhenomPv2; a launch beyond the fixed scratch is refused
before anything is enqueued."""
import ctypes

import numpy as np
import pytest

import gw_analysis_tools_amd as G
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu


def _event():
    f = W.uniform_grid(20., 2048., 0.125)
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    inj = W.injection_params(1)[0]
    p = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item()) for k in inj.dtype.names})
    g = W.rng(1, 0)
    data = [r + W.noise(f, psd, g) for r in G.create_coherent_GW_detection(f, p, dets)]
    return G.Event(f, data, [psd] * 3, dets), inj


@pytest.mark.parametrize("model,prec", [(G.IMRPHENOMD, G.FP64), (G.IMRPHENOMD, G.MIXED),
                                        (G.IMRPHENOMPV2, G.MIXED)])
def test_noalloc_equals_allocating_entry_point(model, prec):
    G.check(G.lib().gwh_init(0))
    lib = G.lib()
    ev, inj = _event()
    P = W.perturbed_params(inj, 512, step=3) if model == G.IMRPHENOMD else W.precessing_params(512)
    d_p = G.DeviceArray.from_host(P)
    a, b = G.DeviceArray(512 * 8), G.DeviceArray(512 * 8)
    for n in (512, 77):  # the second call smaller: stale records of the first must not leak into it
        print("noalloc model %d prec %d n %d" % (model, prec, n), flush=True)
        G.check(lib.gwh_loglike_batch_device(None, ev.id, model, prec, d_p.ptr, n, a.ptr, None))
        G.check(lib.gwh_loglike_batch_device_noalloc(None, ev.id, model, prec, d_p.ptr, n, b.ptr, None))
        G.synchronize()
        x, y = a.to_host(np.float64, n), b.to_host(np.float64, n)
        assert np.all(np.isfinite(x))
        assert np.array_equal(x, y), np.max(np.abs(x - y))
    for d in (d_p, a, b):
        d.free()
    ev.close()


def test_noalloc_refuses_beyond_its_scratch():
    G.check(G.lib().gwh_init(0))
    lib = G.lib()
    ev, inj = _event()
    n = 65535  # records alone exceed the fixed array
    rc = lib.gwh_loglike_batch_device_noalloc(None, ev.id, G.IMRPHENOMD, G.FP64, ctypes.c_void_p(0), n,
                                              ctypes.c_void_p(0), None)
    assert rc == G.GWH_ERR_NOMEM
    assert b"allocation-free scratch" in lib.gwh_last_error()
    ev.close()
