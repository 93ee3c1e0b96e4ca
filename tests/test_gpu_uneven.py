"""Uneven multi-event batches (events from ~1e3 to ~1e5 bins in one launch through the flattened
(instance, tile) work lists): FP64 equals each event's own single-event batch bit for bit and the
oracle within 1e-9; mixed within 1e-3 of FP64; the autotuner keeps batch results reproducible."""
import ctypes

import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu


def _events(n_events=16, chains=3):
    specs = W.config2(n_events=n_events, chains=chains)
    evs, P, ids = [], [], []
    g = W.rng(2, 1)
    for k, sp in enumerate(specs):
        f, psd = sp["f"], sp["psd"]
        inj = sp["injection"]
        pinj = G.make_params(**{n: (list(inj[n]) if inj[n].shape else inj[n].item())
                                for n in inj.dtype.names})
        resp = G.create_coherent_GW_detection(f, pinj, sp["dets"])
        data = [r + W.noise(f, psd, g) for r in resp]
        e = G.Event(f, data, [psd] * 3, sp["dets"])
        evs.append((e, sp, data))
        P.append(sp["params"])
        ids += [e.id] * len(sp["params"])
    return evs, np.concatenate(P), np.array(ids, dtype=np.int32)


def _oracle_params(rec):
    return O.make_params(**{k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names})


def test_uneven_batch_equals_single_event_batches_and_oracle():
    evs, P, ids = _events()
    # interleave instances of different events
    order = np.random.default_rng(3).permutation(len(P))
    P, ids = P[order], ids[order]
    full = G.loglike_batch(ids[0], P, G.IMRPHENOMD, G.FP64, events=list(ids))
    byid = {e.id: (e, sp, data) for e, sp, data in evs}
    for i in range(len(P)):
        e = byid[ids[i]][0]
        alone = G.loglike_batch(e, P[i:i + 1], G.IMRPHENOMD, G.FP64)[0]
        assert full[i] == alone, (i, full[i], alone)
    for i in range(0, len(P), 7):
        e, sp, data = byid[ids[i]]
        ref = O.network_loglike(sp["f"], data, [sp["psd"]] * 3, sp["dets"], _oracle_params(P[i]))
        assert abs(full[i] - ref) <= 1e-9 * max(abs(ref), 1.0), (i, full[i], ref)
    mixed = G.loglike_batch(ids[0], P, G.IMRPHENOMD, G.MIXED, events=list(ids))
    d = np.abs(mixed - full)
    print("uneven batch: %d instances, N %d..%d, mixed vs FP64 max %.3e" %
          (len(P), min(len(s["f"]) for _, s, _ in evs), max(len(s["f"]) for _, s, _ in evs), d.max()))
    assert d.max() <= 1e-3
    for e, _, _ in evs:
        e.close()


def test_autotuner_table_and_reproducibility():
    evs, P, ids = _events(n_events=10, chains=2)
    d_params = G.DeviceArray.from_host(P)
    d_ids = G.DeviceArray.from_host(ids)
    d_out = G.DeviceArray(len(P) * 8)
    table = np.zeros(32, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    G.check(G.lib().gwh_autotune_bins_per_lane(ids.ctypes.data_as(ip), d_ids.ptr, G.IMRPHENOMD, G.MIXED,
                                               d_params.ptr, len(P), d_out.ptr, 3, table.ctypes.data_as(ip)))
    present = {int(np.floor(np.log2(len(sp["f"])))) for _, sp, _ in evs}
    assert {i for i in range(32) if table[i]} == present
    assert set(table[table > 0]) <= {1, 2, 4, 8}
    print("autotuned bins per lane by log2 N:", {i: int(table[i]) for i in range(32) if table[i]})
    # after tuning, batch and single-event results still agree bit for bit
    full = G.loglike_batch(ids[0], P, G.IMRPHENOMD, G.MIXED, events=list(ids))
    G.check(G.lib().gwh_loglike_multi_device(ids.ctypes.data_as(ip), d_ids.ptr, G.IMRPHENOMD, G.MIXED, d_params.ptr,
                                             len(P), d_out.ptr, None))
    G.synchronize()
    assert np.array_equal(d_out.to_host(np.float64, len(P)), full)
    for i in range(len(P)):
        assert G.loglike_batch(int(ids[i]), P[i:i + 1], G.IMRPHENOMD, G.MIXED)[0] == full[i]
    G.set_bins_per_lane(0)
    for e, _, _ in evs:
        e.close()
