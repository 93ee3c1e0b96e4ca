"""Thin Python view of the C ABI (include/gwat_hip.h), used by tests and bench.py.

Names mirror the reference's entry points (src/waveform_generator.cpp
fourier_waveform, src/waveform_util.cpp fourier_detector_response,
src/mcmc_gw.cpp Log_Likelihood / MCMC_likelihood_extrinsic, src/fisher.cpp
fisher_numerical); all compute happens in libgwat_hip.so on the GPU.
"""
import ctypes
import os

import numpy as np

from . import GWHError, Params, PTConfig, check, lib, LIB_PATH, PARAMS_DTYPE, IMRPHENOMD, FP64, BASIS_PHENOMD_MCMC

__all__ = ["fourier_waveform", "fourier_detector_response", "create_coherent_GW_detection", "work_counts",
           "log_likelihood", "log_likelihood_internal", "mcmc_likelihood_extrinsic", "fisher_numerical",
           "fisher_derivatives", "unpack_parameters", "repack_parameters", "detector_id", "dtoa",
           "gps_to_gmst_radian", "calculate_snr_internal", "data_snr", "snr_batch", "Broker", "Event", "loglike_batch",
           "fisher_batch", "DeviceArray", "set_bins_per_lane", "set_tile_lanes", "autotune_tiles", "synchronize", "LoglikeGraph", "DevicePT",
           "maximized_log_likelihood_aligned_internal", "maximized_log_likelihood_unaligned_internal",
           "maximized_coal_log_likelihood_internal", "maximized_loglike_batch", "gauleg", "THREADED_CALLS",
           "write_threaded_calls_input"]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _interleave(z):
    z = np.asarray(z, dtype=np.complex128)
    return np.ascontiguousarray(z).view(np.float64)


def _cplx(x):
    return x.view(np.complex128)


def synchronize():
    check(lib().gwh_synchronize())


def set_bins_per_lane(bpl):
    check(lib().gwh_set_bins_per_lane(int(bpl)))


def set_tile_lanes(lanes):
    """lanes (threads) per likelihood tile: 64, 128, 256, 512 or 1024; 0 = automatic (autotuned, else 256)"""
    check(lib().gwh_set_tile_lanes(int(lanes)))


def autotune_tiles(event_ids, d_event_ids, model, precision, d_params, d_logl, reps=3):
    """Tile the events of a batch by timing it over lanes per tile x bins per lane per event-size bucket
    (gwh_autotune_tiles).  event_ids: the batch's host ids (int32, one per instance); d_event_ids, d_params,
    d_logl: DeviceArrays of the same batch.  Returns {floor(log2 N): (bins_per_lane, lanes)}."""
    ids = np.ascontiguousarray(event_ids, dtype=np.int32)
    bpl, lanes = np.zeros(32, dtype=np.int32), np.zeros(32, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    check(lib().gwh_autotune_tiles(ids.ctypes.data_as(ip), d_event_ids.ptr, int(model), int(precision),
                                   d_params.ptr, len(ids), d_logl.ptr, int(reps), bpl.ctypes.data_as(ip),
                                   lanes.ctypes.data_as(ip)))
    return {i: (int(bpl[i]), int(lanes[i])) for i in range(32) if bpl[i]}


def fourier_waveform(f, p, model=IMRPHENOMD):
    f = _f64(f)
    n = len(f)
    hp = np.zeros(2 * n)
    hc = np.zeros(2 * n)
    check(lib().gwh_fourier_waveform(_dp(f), n, _dp(hp), _dp(hc), model, ctypes.byref(p)))
    return _cplx(hp), _cplx(hc)


def fourier_detector_response(f, p, det, model=IMRPHENOMD):
    f = _f64(f)
    r = np.zeros(2 * len(f))
    check(lib().gwh_fourier_detector_response(_dp(f), len(f), _dp(r), det, model, ctypes.byref(p)))
    return _cplx(r)


def create_coherent_GW_detection(f, p, dets, model=IMRPHENOMD):
    f = _f64(f)
    D = len(dets)
    out = np.zeros(2 * D * len(f))
    di = (ctypes.c_int * D)(*dets)
    check(lib().gwh_create_coherent_GW_detection(di, D, _dp(f), len(f), model, ctypes.byref(p), _dp(out)))
    return [_cplx(out[2 * len(f) * d:2 * len(f) * (d + 1)]) for d in range(D)]


NREGION_COUNTS = 10


def work_counts(event, params, model=IMRPHENOMD):
    """bins of every instance by (amplitude region a, phase region p): [n, 3, 3] non-zero bins and [n] zero bins"""
    arr = _as_params_array(params)
    c = np.zeros(NREGION_COUNTS * len(arr), dtype=np.int32)
    ev = event.id if hasattr(event, "id") else int(event)
    check(lib().gwh_work_counts(ev, model, arr.ctypes.data_as(ctypes.c_void_p), len(arr),
                                c.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    c = c.reshape(-1, NREGION_COUNTS)
    return c[:, :9].reshape(-1, 3, 3), c[:, 9]


def _nan_check(v):
    if np.isnan(v):
        msg = lib().gwh_last_error().decode(errors="replace")
        if msg:
            raise GWHError(msg)
    return v


def log_likelihood(data, psd, f, p, det, model=IMRPHENOMD):
    d = _interleave(data)
    s = _f64(psd)
    f = _f64(f)
    return lib().gwh_log_likelihood(_dp(d), _dp(s), _dp(f), len(f), ctypes.byref(p), det, model)


def log_likelihood_internal(data, psd, f, response, weights=None, log10F=False, method=0):
    d = _interleave(data)
    r = _interleave(response)
    s = _f64(psd)
    f = _f64(f)
    w = _f64(weights) if weights is not None else None
    return lib().gwh_log_likelihood_internal(_dp(d), _dp(s), _dp(f), _dp(w) if w is not None else None, _dp(r),
                                             len(f), int(log10F), method)


def mcmc_likelihood_extrinsic(f, data, psd, dets, p, model=IMRPHENOMD, T_eff=0.0, save_waveform=True, weights=None,
                              log10F=False, method=0):
    """weights (Gauss-Legendre, method GAUSSLEG) are shared by every detector, as the detectors share one grid"""
    f = _f64(f)
    n = len(f)
    D = len(dets)
    dd = [_interleave(x) for x in data]
    ss = [_f64(x) for x in psd]
    DP = ctypes.POINTER(ctypes.c_double)
    fl = (DP * D)(*[_dp(f)] * D)
    dl = (DP * D)(*[_dp(x) for x in dd])
    sl = (DP * D)(*[_dp(x) for x in ss])
    w = _f64(weights) if weights is not None else None
    wl = (DP * D)(*[_dp(w)] * D) if w is not None else None
    lens = (ctypes.c_int * D)(*[n] * D)
    di = (ctypes.c_int * D)(*dets)
    return lib().gwh_mcmc_likelihood_extrinsic(int(save_waveform), ctypes.byref(p), model, lens, fl, dl, sl, wl,
                                               method, int(log10F), di, D, T_eff)


def fisher_numerical(f, psd, basis, det, ref_det, dim, p, order=2, epsilon=1e-8, model=None):
    """model None: the GR model of the basis; a ppE model appends p.ppe_nmod betas to the basis, a gIMR model the
    p.gimr_nlist listed deltas.  epsilon > 0 is the absolute step of fisher.cpp:158; epsilon < 0 a step of
    |epsilon| * |theta_i| for each basis coordinate (|epsilon| where theta_i is 0)"""
    f = _f64(f)
    s = _f64(psd)
    out = np.zeros(dim * dim)
    if model is None:
        check(lib().gwh_fisher_numerical(_dp(f), len(f), basis, det, ref_det, _dp(out), dim, ctypes.byref(p), order,
                                         _dp(s), epsilon))
    else:
        check(lib().gwh_fisher_numerical_model(_dp(f), len(f), model, basis, det, ref_det, _dp(out), dim,
                                               ctypes.byref(p), order, _dp(s), epsilon))
    return out.reshape(dim, dim)


def fisher_derivatives(f, basis, det, ref_det, dim, p, order=2, epsilon=1e-8, model=None):
    """bare derivatives dh/dtheta_i (dim x len(f) complex), as calculate_derivatives returns them"""
    f = _f64(f)
    out = np.zeros(2 * dim * len(f))
    if model is None:
        check(lib().gwh_fisher_derivatives(_dp(f), len(f), basis, det, ref_det, dim, ctypes.byref(p), order, epsilon,
                                           _dp(out)))
    else:
        check(lib().gwh_fisher_derivatives_model(_dp(f), len(f), model, basis, det, ref_det, dim, ctypes.byref(p),
                                                 order, epsilon, _dp(out)))
    return _cplx(out).reshape(dim, len(f))


def unpack_parameters(p, basis, dim):
    v = np.zeros(dim)
    lf = (ctypes.c_int * dim)()
    check(lib().gwh_unpack_parameters(ctypes.byref(p), basis, _dp(v), lf))
    return v, np.array(list(lf), dtype=bool)


def repack_parameters(v, basis, template):
    out = Params()
    ctypes.pointer(out)[0] = template
    v = _f64(v)
    check(lib().gwh_repack_parameters(_dp(v), basis, ctypes.byref(out)))
    return out


def gauleg(x1, x2, n):
    """Gauss-Legendre nodes (ascending) and weights of the n-point rule on [x1, x2] (gauleg)"""
    x, w = np.zeros(n), np.zeros(n)
    check(lib().gwh_gauleg(x1, x2, _dp(x), _dp(w), n))
    return x, w


def detector_id(name):
    return lib().gwh_detector_id(name.encode())


def dtoa(ra, dec, gmst, det1, det2):
    return _nan_check(lib().gwh_dtoa(ra, dec, gmst, det1, det2))


def gps_to_gmst_radian(gps):
    return lib().gwh_gps_to_gmst_radian(gps)


def calculate_snr_internal(psd, waveform, f, weights=None, log10F=False, method=0):
    w = _f64(weights) if weights is not None else None
    return _nan_check(lib().gwh_calculate_snr_internal(_dp(_f64(psd)), _dp(_interleave(waveform)), _dp(_f64(f)),
                                                       len(f), method, _dp(w) if w is not None else None,
                                                       int(log10F)))


def data_snr(f, data, response, psd):
    return _nan_check(lib().gwh_data_snr(_dp(_f64(f)), len(f), _dp(_interleave(data)), _dp(_interleave(response)),
                                         _dp(_f64(psd))))


def snr_batch(event, params, model=IMRPHENOMD, precision=FP64, events=None):
    arr = _as_params_array(params)
    n = len(arr)
    out = np.zeros(n)
    ev0 = event.id if isinstance(event, Event) else int(event)
    evs = None
    if events is not None:
        e = np.ascontiguousarray([x.id if isinstance(x, Event) else int(x) for x in events], dtype=np.int32)
        evs = e.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    check(lib().gwh_snr_batch(evs, ev0, model, precision, arr.ctypes.data_as(ctypes.c_void_p), n, _dp(out)))
    return out


class Event:
    """A resident event: shared grid, per-detector data and PSD kept in HBM."""

    def __init__(self, f, data, psd, dets, weights=None, log10F=False):
        """weights: Gauss-Legendre weights of the grid (GAUSSLEG rule; None: Simpson's rule)"""
        f = _f64(f)
        D = len(dets)
        dat = np.concatenate([_interleave(x) for x in data])
        ps = np.concatenate([_f64(x) for x in psd])
        ev = ctypes.c_int(-1)
        di = (ctypes.c_int * D)(*dets)
        if weights is None and not log10F:
            check(lib().gwh_event_create(_dp(f), len(f), D, di, _dp(dat), _dp(ps), ctypes.byref(ev)))
        else:
            w = _f64(weights) if weights is not None else None
            check(lib().gwh_event_create_weighted(_dp(f), len(f), D, di, _dp(dat), _dp(ps), 1 if w is not None else 0,
                                                  _dp(w) if w is not None else None, int(log10F), ctypes.byref(ev)))
        self.id = ev.value
        self.n = len(f)
        self.ndet = D

    @property
    def nbytes(self):
        return lib().gwh_event_bytes(self.id)

    def close(self):
        if self.id >= 0:
            check(lib().gwh_event_destroy(self.id))
            self.id = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _as_params_array(params):
    if isinstance(params, np.ndarray) and params.dtype == PARAMS_DTYPE:
        return np.ascontiguousarray(params)
    arr = np.zeros(len(params), dtype=PARAMS_DTYPE)
    for i, p in enumerate(params):
        arr[i] = np.frombuffer(bytes(p), dtype=PARAMS_DTYPE)[0]
    return arr


def loglike_batch(event, params, model=IMRPHENOMD, precision=FP64, events=None):
    arr = _as_params_array(params)
    n = len(arr)
    out = np.zeros(n)
    ev0 = event.id if isinstance(event, Event) else int(event)
    evs = None
    if events is not None:
        e = np.ascontiguousarray([x.id if isinstance(x, Event) else int(x) for x in events], dtype=np.int32)
        evs = e.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    check(lib().gwh_loglike_batch(evs, ev0, model, precision, arr.ctypes.data_as(ctypes.c_void_p), n, _dp(out)))
    return out


def fisher_batch(event, params, model, basis, dim, precision=FP64, epsilon=1e-8, with_sigma=False):
    arr = _as_params_array(params)
    n = len(arr)
    out = np.zeros(n * dim * dim)
    sig = np.zeros(n * dim) if with_sigma else None
    ev0 = event.id if isinstance(event, Event) else int(event)
    check(lib().gwh_fisher_batch(ev0, model, basis, precision, arr.ctypes.data_as(ctypes.c_void_p), n, dim, epsilon,
                                 _dp(out), _dp(sig) if sig is not None else None))
    F = out.reshape(n, dim, dim)
    return (F, sig.reshape(n, dim)) if with_sigma else F


class DeviceArray:
    """Raw device allocation owned by the engine's allocator (bench/test plumbing)."""

    def __init__(self, nbytes):
        self.ptr = lib().gwh_device_alloc(nbytes)
        if not self.ptr:
            raise GWHError("device allocation failed: " + lib().gwh_last_error().decode())
        self.nbytes = nbytes

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.nbytes)
        check(lib().gwh_memcpy_h2d(d.ptr, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes))
        return d

    def to_host(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        check(lib().gwh_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            check(lib().gwh_device_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Broker:
    """Continuous-batching request broker (gwh_broker_*): submit from any thread, wait on tickets."""

    def __init__(self, hbm_budget_bytes=200 << 30, max_batch_instances=8192):
        b = ctypes.c_int(-1)
        check(lib().gwh_broker_create(int(hbm_budget_bytes), int(max_batch_instances), ctypes.byref(b)))
        self.id = b.value
        self._keep = {}

    def submit_loglike(self, events, params, model=IMRPHENOMD, precision=FP64):
        """events: one event (Event or id) or one per instance; returns (ticket, output array)"""
        arr = _as_params_array(params)
        n = len(arr)
        out = np.zeros(n)
        if isinstance(events, (list, tuple, np.ndarray)):
            ev = np.ascontiguousarray([x.id if isinstance(x, Event) else int(x) for x in events], dtype=np.int32)
            evp, ev0 = ev.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(ev[0])
        else:
            ev, evp, ev0 = None, None, (events.id if isinstance(events, Event) else int(events))
        t = ctypes.c_longlong(0)
        check(lib().gwh_broker_submit_loglike(self.id, evp, ev0, model, precision, arr.ctypes.data_as(ctypes.c_void_p),
                                              n, _dp(out), ctypes.byref(t)))
        self._keep[t.value] = (arr, ev, out)
        return t.value, out

    def submit_fisher(self, event, params, model, basis, dim, precision=FP64, epsilon=1e-8, with_sigma=False):
        arr = _as_params_array(params)
        n = len(arr)
        out = np.zeros(n * dim * dim)
        sig = np.zeros(n * dim) if with_sigma else None
        ev0 = event.id if isinstance(event, Event) else int(event)
        t = ctypes.c_longlong(0)
        check(lib().gwh_broker_submit_fisher(self.id, ev0, model, basis, precision, arr.ctypes.data_as(ctypes.c_void_p),
                                             n, dim, epsilon, _dp(out), _dp(sig) if sig is not None else None,
                                             ctypes.byref(t)))
        self._keep[t.value] = (arr, out, sig)
        return t.value, out.reshape(n, dim, dim), (sig.reshape(n, dim) if sig is not None else None)

    def wait(self, ticket, timeout_ms=-1.0):
        """waits for the ticket; returns the latency in microseconds (submission to completion)"""
        lat = np.zeros(1)
        st = lib().gwh_broker_wait(self.id, ticket, timeout_ms, _dp(lat))
        self._keep.pop(ticket, None)
        check(st)
        return float(lat[0])

    def stats(self):
        s = np.zeros(9)
        check(lib().gwh_broker_stats(self.id, _dp(s)))
        keys = ("batches", "instances", "fisher_requests", "max_requests_per_batch", "admission_waits",
                "peak_device_bytes", "tickets_outstanding", "fisher_batches", "max_fisher_requests_per_batch")
        return {k: (int(v) if k != "peak_device_bytes" else int(v)) for k, v in zip(keys, s)}

    def close(self):
        if self.id >= 0:
            check(lib().gwh_broker_destroy(self.id))
            self.id = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LoglikeGraph:
    """gwh_loglike_graph_*: the batched likelihood on fixed device buffers, captured once and replayed."""

    def __init__(self, event, d_params, n, d_logl, model=IMRPHENOMD, precision=FP64, events=None, d_events=None):
        g = ctypes.c_int(-1)
        if events is not None:
            self._ev = np.ascontiguousarray([x.id if isinstance(x, Event) else int(x) for x in events], dtype=np.int32)
            evp, ev0 = self._ev.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(self._ev[0])
        else:
            evp, ev0 = None, (event.id if isinstance(event, Event) else int(event))
        check(lib().gwh_loglike_graph_create(evp, d_events, ev0, model, precision, d_params, n, d_logl,
                                             ctypes.byref(g)))
        self.id = g.value

    def launch(self, stream=None):
        check(lib().gwh_graph_launch(self.id, stream))

    def synchronize(self):
        check(lib().gwh_graph_synchronize(self.id))

    def close(self):
        if self.id >= 0:
            check(lib().gwh_graph_destroy(self.id))
            self.id = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DevicePT:
    """gwh_pt_*: parallel tempering with the whole step (proposal, prior, likelihoods, Metropolis-Hastings,
    ladder swaps) on the device, one graph replay per step.  Instance i = temp * n_chains + chain."""

    def __init__(self, event, template, theta0, widths, temps, n_chains, lower, upper, model=IMRPHENOMD,
                 precision=FP64, basis=BASIS_PHENOMD_MCMC, seed=20261203, swap_every=10, guard=True, tol=1e-3,
                 steps_per_graph=1):
        temps = np.ascontiguousarray(temps, dtype=np.float64)
        self.nt, self.nc = len(temps), int(n_chains)
        self.n = self.nt * self.nc
        self.dim = len(lower)
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(theta0, dtype=np.float64), (self.n, self.dim)))
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(widths, dtype=np.float64), (self.n, self.dim)))
        lo = np.ascontiguousarray(lower, dtype=np.float64)
        hi = np.ascontiguousarray(upper, dtype=np.float64)
        if isinstance(template, np.void):
            tmpl = np.array([template], dtype=PARAMS_DTYPE)
        else:
            tmpl = _as_params_array(template)[:1].copy()
        cfg = PTConfig(n_temps=self.nt, n_chains=self.nc, model=model, precision=precision, basis=basis,
                       swap_every=int(swap_every), guard=int(bool(guard)), steps_per_graph=int(steps_per_graph),
                       tol=float(tol), seed=int(seed))
        h = ctypes.c_int(-1)
        check(lib().gwh_pt_create(event.id if isinstance(event, Event) else int(event), ctypes.byref(cfg),
                                  _dp(temps), _dp(w), _dp(lo), _dp(hi),
                                  tmpl.ctypes.data_as(ctypes.POINTER(Params)), _dp(th), ctypes.byref(h)))
        self.id = h.value
        self.steps_per_graph = int(steps_per_graph)

    def run(self, steps, sync=True):
        check(lib().gwh_pt_run(self.id, int(steps)))
        if sync:
            check(lib().gwh_pt_synchronize(self.id))

    def synchronize(self):
        check(lib().gwh_pt_synchronize(self.id))

    def state(self):
        th = np.zeros((self.n, self.dim))
        ll = np.zeros(self.n)
        st = ctypes.c_longlong(0)
        check(lib().gwh_pt_state(self.id, _dp(th), _dp(ll), ctypes.byref(st)))
        return th, ll, st.value

    def last_step(self):
        prop = np.zeros((self.n, self.dim))
        llp = np.zeros(self.n)
        acc = np.zeros(self.n, dtype=np.int32)
        sw = np.zeros(max(self.nt - 1, 0) * self.nc, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        check(lib().gwh_pt_last_step(self.id, _dp(prop), _dp(llp), acc.ctypes.data_as(ip), sw.ctypes.data_as(ip)))
        return prop, llp, acc.astype(bool), sw.reshape(max(self.nt - 1, 0), self.nc).astype(bool)

    def counters(self):
        acc = np.zeros(self.n, dtype=np.int64)
        sw = np.zeros(max(self.nt - 1, 0) * self.nc, dtype=np.int64)
        g = ctypes.c_longlong(0)
        lp = ctypes.POINTER(ctypes.c_longlong)
        check(lib().gwh_pt_counters(self.id, acc.ctypes.data_as(lp), sw.ctypes.data_as(lp), ctypes.byref(g)))
        return acc, sw.reshape(max(self.nt - 1, 0), self.nc), g.value

    def close(self):
        if self.id >= 0:
            check(lib().gwh_pt_destroy(self.id))
            self.id = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def maximized_log_likelihood_aligned_internal(data, psd, f, response):
    """maximized_Log_Likelihood_aligned_spin_internal (src/mcmc_gw.cpp:595-654)"""
    f = _f64(f)
    v = lib().gwh_maximized_log_likelihood_aligned_internal(_dp(_interleave(data)), _dp(_f64(psd)), _dp(f),
                                                            _dp(_interleave(response)), len(f))
    return _nan_check(v)


def maximized_log_likelihood_unaligned_internal(data, psd, f, hplus, hcross):
    """maximized_Log_Likelihood_unaligned_spin_internal (src/mcmc_gw.cpp:660-800)"""
    f = _f64(f)
    v = lib().gwh_maximized_log_likelihood_unaligned_internal(_dp(_interleave(data)), _dp(_f64(psd)), _dp(f),
                                                              _dp(_interleave(hplus)), _dp(_interleave(hcross)), len(f))
    return _nan_check(v)


def maximized_coal_log_likelihood_internal(data, psd, f, response):
    """maximized_coal_Log_Likelihood_internal (src/mcmc_gw.cpp:493-556): (logL, tc, phiRef)"""
    f = _f64(f)
    tc, ph = np.zeros(1), np.zeros(1)
    v = lib().gwh_maximized_coal_log_likelihood_internal(_dp(_interleave(data)), _dp(_f64(psd)), _dp(f),
                                                         _dp(_interleave(response)), len(f), _dp(tc), _dp(ph))
    return _nan_check(v), float(tc[0]), float(ph[0])


def maximized_loglike_batch(event, params, model=IMRPHENOMD):
    """the intrinsic MCMC likelihood of every parameter set on a resident event (src/mcmc_gw.cpp:2607-2700)"""
    arr = _as_params_array(params)
    out = np.zeros(len(arr))
    ev0 = event.id if isinstance(event, Event) else int(event)
    check(lib().gwh_maximized_loglike_batch(ev0, model, arr.ctypes.data_as(ctypes.c_void_p), len(arr), _dp(out)))
    return out


THREADED_CALLS = os.path.join(os.path.dirname(LIB_PATH), "gwh_threaded_calls")


def write_threaded_calls_input(path, f, data, psd, dets, params):
    """Input file of the threaded single-call driver (host/gwh_threaded_calls.cpp): int32 n, ndet, det[ndet],
    count, then f[n], per detector data[n] (complex) and psd[n], then `count` gwh_params records."""
    f = _f64(f)
    P = _as_params_array(params)
    with open(path, "wb") as fh:
        np.array([len(f), len(dets)] + list(dets) + [len(P)], dtype=np.int32).tofile(fh)
        f.tofile(fh)
        for d in range(len(dets)):
            np.ascontiguousarray(data[d], dtype=np.complex128).tofile(fh)
            _f64(psd[d]).tofile(fh)
        P.tofile(fh)
