"""Tilings of the per-bin likelihood kernels: every lanes-per-tile (64..1024) x bins-per-lane (1..8) class gives the
oracle's FP64 log-likelihood within 1e-9 (relative) on the c1 network grid, the tilings agree with each other to
summation order (1e-11), and mixed stays within 1e-3 of FP64; short and boundary grid lengths (2 .. 8193 bins,
where a tile holds the whole grid or one bin past a frequency-minimum chunk) match the oracle in the extreme
tilings; the tiling autotuner returns valid classes, and after it a multi-event batch whose events are tiled
differently equals each event's own batch bit for bit."""
import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu

LANES = (64, 128, 256, 512, 1024)
BPL = (1, 2, 4, 8)


def _oracle_params(rec):
    return O.make_params(**{k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names})


def _c1_event():
    cfg = W.config1()
    f, psd, dets = cfg["f"], cfg["psd"], cfg["dets"]
    inj = W.injection_params(1)[0]
    pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item()) for k in inj.dtype.names})
    data = [r + W.noise(f, psd, W.rng(1, 0)) for r in G.create_coherent_GW_detection(f, pinj, dets)]
    return G.Event(f, data, [psd] * 3, dets), f, data, psd, dets, inj


@pytest.fixture(autouse=True)
def _default_tiling():
    yield
    G.set_tile_lanes(0)
    G.set_bins_per_lane(0)


def test_every_tiling_class_matches_the_oracle():
    ev, f, data, psd, dets, inj = _c1_event()
    # perturbed chains plus heavy binaries whose 0.2/M cutoff falls inside the grid (tiles skipped)
    P = W.perturbed_params(inj, 24, step=5)
    heavy = P[:4].copy()
    heavy["mass1"] = [150.0, 210.0, 260.0, 320.0]
    heavy["mass2"] = [120.0, 190.0, 240.0, 300.0]
    P = np.concatenate([P, heavy])
    ref = np.array([O.network_loglike(f, data, [psd] * 3, dets, _oracle_params(P[i])) for i in range(len(P))])
    base = None
    for lanes in LANES:
        G.set_tile_lanes(lanes)
        for b in BPL:
            G.set_bins_per_lane(b)
            fp64 = G.loglike_batch(ev, P, G.IMRPHENOMD, G.FP64)
            mixed = G.loglike_batch(ev, P, G.IMRPHENOMD, G.MIXED)
            err = np.abs(fp64 - ref) / np.maximum(np.abs(ref), 1.0)
            assert err.max() <= 1e-9, (lanes, b, err.max())
            assert np.abs(mixed - fp64).max() <= 1e-3, (lanes, b)
            if base is None:
                base = fp64
            assert np.abs(fp64 - base).max() <= 1e-11 * np.abs(base).max(), (lanes, b)
    print("tilings: 20 classes within %.1e of the oracle" % err.max())
    ev.close()


def test_pv2_tilings_agree():
    ev, f, data, psd, dets, inj = _c1_event()
    P = W.precessing_params(16)
    out = {}
    for lanes, b in ((64, 1), (1024, 8), (128, 4), (512, 2)):
        G.set_tile_lanes(lanes)
        G.set_bins_per_lane(b)
        out[(lanes, b)] = (G.loglike_batch(ev, P, G.IMRPHENOMPV2, G.FP64), G.loglike_batch(ev, P, G.IMRPHENOMPV2,
                                                                                            G.MIXED))
    ref = np.array([O.network_loglike(f, data, [psd] * 3, dets, _oracle_params(P[i]), model=O.IMRPHENOMPV2)
                    for i in range(0, len(P), 5)])
    for (lanes, b), (fp64, mixed) in out.items():
        assert np.allclose(fp64[::5], ref, rtol=1e-9, atol=1e-9), (lanes, b)
        assert np.abs(mixed - fp64).max() <= 1e-3, (lanes, b)
    ev.close()


@pytest.mark.parametrize("N", [2, 63, 64, 65, 1000, 8191, 8193])
def test_short_and_boundary_grids(N):
    f = 20.0 + 0.25 * np.arange(N)
    psd = W.analytic_psd(f)
    dets = [0, 1]
    inj = W.injection_params(1)[0]
    pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item()) for k in inj.dtype.names})
    data = [r + 0.1 * W.noise(f, psd, W.rng(N, 3)) for r in G.create_coherent_GW_detection(f, pinj, dets)]
    ev = G.Event(f, data, [psd] * 2, dets)
    P = W.perturbed_params(inj, 6, step=2)
    ref = np.array([O.network_loglike(f, data, [psd] * 2, dets, _oracle_params(P[i])) for i in range(len(P))])
    for lanes, b in ((64, 1), (1024, 8), (256, 2)):
        G.set_tile_lanes(lanes)
        G.set_bins_per_lane(b)
        got = G.loglike_batch(ev, P, G.IMRPHENOMD, G.FP64)
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(ref).max())), (N, lanes, b, got, ref)
    ev.close()


def test_tiling_autotuner_and_mixed_tilings_in_one_batch():
    specs = W.config2(n_events=8, chains=2)
    evs, P, ids = [], [], []
    g = W.rng(2, 1)
    for sp in specs:
        inj = sp["injection"]
        pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item()) for k in inj.dtype.names})
        data = [r + W.noise(sp["f"], sp["psd"], g) for r in G.create_coherent_GW_detection(sp["f"], pinj, sp["dets"])]
        evs.append((G.Event(sp["f"], data, [sp["psd"]] * 3, sp["dets"]), sp, data))
        P.append(sp["params"])
        ids += [evs[-1][0].id] * len(sp["params"])
    P, ids = np.concatenate(P), np.array(ids, dtype=np.int32)
    d_params, d_ids, d_out = G.DeviceArray.from_host(P), G.DeviceArray.from_host(ids), G.DeviceArray(len(P) * 8)
    choice = G.autotune_tiles(ids, d_ids, G.IMRPHENOMD, G.MIXED, d_params, d_out, reps=2)
    present = {int(np.floor(np.log2(len(sp["f"])))) for _, sp, _ in evs}
    assert set(choice) == present
    assert all(b in BPL and l in LANES for b, l in choice.values()), choice
    print("autotuned tiling by log2 N:", choice)
    full = G.loglike_batch(ids[0], P, G.IMRPHENOMD, G.FP64, events=list(ids))
    byid = {e.id: (e, sp, data) for e, sp, data in evs}
    for i in range(len(P)):
        assert G.loglike_batch(int(ids[i]), P[i:i + 1], G.IMRPHENOMD, G.FP64)[0] == full[i], i
    for i in range(0, len(P), 5):
        e, sp, data = byid[ids[i]]
        ref = O.network_loglike(sp["f"], data, [sp["psd"]] * 3, sp["dets"], _oracle_params(P[i]))
        assert abs(full[i] - ref) <= 1e-9 * max(abs(ref), 1.0), (i, full[i], ref)
    for d in (d_params, d_ids, d_out):
        d.free()
    for e, _, _ in evs:
        e.close()
