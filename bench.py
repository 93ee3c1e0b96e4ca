#!/usr/bin/env python3
"""Benchmark: IMRPhenomD 3-detector network log-likelihood evaluations per second on one MI355X.

Workload (the batched parallel-tempering ensemble, 16225 bins 20-2048 Hz at df = 1/8 Hz, Hanford /
Livingston / Virgo, 16 temperatures x 32 chains = 512 instances per step): one step evaluates the
network log-likelihood of 512 parameter vectors scattered around the injection.  All parameter batches
and the event (data, PSD, weights) are resident in HBM before the timed region; every step is one
enqueue of the batched engine followed by a stream synchronisation (a sampler reads its logL every
step), so per-step latencies give p50/p99.

`value` is the mixed-precision batched engine (FP64 except the merger-ringdown exponential, in FP32; sin/cos
of the detector phases through the FP64 tables), on the tiling the autotuner picks for the event before the
warm-up.  Reported beside it: the FP64 batched engine, the
plain FP64 one-instance-per-launch path, max |delta logL| mixed vs FP64 over every timed evaluation, the same for
IMRPhenomPv2 on the same grid (1024 chains; its mixed mode keeps FP64 table-driven sin/cos and exponentials),
kernel-level roofline from HIP events, the single-detector workload with its sky-averaged Fisher, batched
13-parameter IMRPhenomPv2 Fisher matrices, the tempering step on the device, the FFT-maximised likelihood, the
uneven multi-event batch, the serving broker, the reference-signature C++ layer called from many host threads,
and the CPU oracle (the reference algorithm restated in C, one instance per call, OpenMP over instances) timed
on bounded samples on the host cores.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "IMRPhenomD 3-det logL evals/s, 1 GPU, mixed vs FP64 at |ΔlogL|≤1e-3; p99 ms"
PEAK_VALU_TFLOPS = 78.6   # MI355X vector FMA peak without packed FP32 (FP64 and non-packed FP32 issue alike)
PEAK_HBM_GBS = 8000.0
# FLOP-equivalents (fe) per non-zero bin, frozen from the operation count of the reference algorithm as the
# oracle restates it (oracle/build/opcount; add/mul/cmp 1, div/sqrt 4, exp/log 16, pow 32, sin+cos 24, atan 20):
# a bin in amplitude region a and phase region p costs FE_COMMON + FE_AMP[a] + FE_PHASE[p], plus FE_DET per
# detector; IMRPhenomPv2 adds FE_PV2_BIN per bin and FE_PV2_DET per detector and bin.
FE_COMMON = 77.0
FE_AMP = (31.0, 19.0, 47.0)     # f < 0.014/M, up to the amplitude peak, above
FE_PHASE = (82.0, 40.0, 57.0)   # f < 0.018/M (with the powers of f it needs), up to f_RD / 2, above
FE_DET = 55.0
FE_PV2_BIN = 302.0
FE_PV2_DET = 0.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fe_per_launch(counts, ndet, pv2=False):
    """counts: (per-instance [3, 3] non-zero bins by amplitude / phase region, zero bins) from work_counts"""
    c, _zero = counts
    per = FE_COMMON + np.add.outer(np.array(FE_AMP), np.array(FE_PHASE))
    nz = c.sum(axis=(1, 2))
    fe = (c * per[None]).sum(axis=(1, 2)) + ndet * nz * FE_DET
    if pv2:
        fe = fe + nz * FE_PV2_BIN + ndet * nz * FE_PV2_DET
    return float(fe.sum()), float(nz.sum())


def bench_uneven(G, W, steps):
    """256 events (total mass 5-200 Msun, durations 4-128 s, 20 Hz to min(2048, 0.3/M) Hz at df = 1/T, H/L/V),
    8 chains each = 2048 instances per launch; the autotuner picks the tiling (lanes per tile x bins per lane) per
    size bucket first."""
    lib = G.lib()
    ip = ctypes.POINTER(ctypes.c_int)
    specs = W.config2()
    g = W.rng(2, 1)
    events, P, ids, nbins = [], [], [], []
    t0 = time.perf_counter()
    for sp in specs:
        f, psd = sp["f"], sp["psd"]
        inj = sp["injection"]
        pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item())
                                for k in inj.dtype.names})
        resp = G.create_coherent_GW_detection(f, pinj, sp["dets"])
        ev = G.Event(f, [r + W.noise(f, psd, g) for r in resp], [psd] * 3, sp["dets"])
        events.append(ev)
        P.append(sp["params"])
        ids += [ev.id] * len(sp["params"])
        nbins += [len(f)] * len(sp["params"])
    t_setup = time.perf_counter() - t0
    P = np.concatenate(P)
    ids = np.array(ids, dtype=np.int32)
    nbins = np.array(nbins)
    n = len(P)
    d_params, d_ids = G.DeviceArray.from_host(P), G.DeviceArray.from_host(ids)
    d_mixed, d_fp64 = G.DeviceArray(n * 8), G.DeviceArray(n * 8)
    resident = sum(ev.nbytes for ev in events)
    t0 = time.perf_counter()
    tiles = G.autotune_tiles(ids, d_ids, G.IMRPHENOMD, G.MIXED, d_params, d_mixed, reps=3)
    t_tune = time.perf_counter() - t0

    def launch(prec, out, idx=None):
        G.check(lib.gwh_loglike_multi_device(ids.ctypes.data_as(ip), d_ids.ptr, G.IMRPHENOMD, prec, d_params.ptr, n,
                                              out.ptr, None))

    for _ in range(3):
        launch(G.MIXED, d_mixed)
        G.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        launch(G.MIXED, d_mixed)
        G.synchronize()
        lat.append(time.perf_counter() - a)
    wall = time.perf_counter() - t0
    launch(G.FP64, d_fp64)
    G.synchronize()
    dmax = float(np.max(np.abs(d_mixed.to_host(np.float64, n) - d_fp64.to_host(np.float64, n))))
    t0 = time.perf_counter()
    for _ in range(max(1, steps // 4)):
        launch(G.FP64, d_fp64)
        G.synchronize()
    fp64_rate = max(1, steps // 4) * n / (time.perf_counter() - t0)
    # per size bucket: the same instances as a batch of their own
    buckets = {}
    lg = np.floor(np.log2(nbins)).astype(int)
    for b in sorted(set(lg)):
        sel = np.nonzero(lg == b)[0]
        Pb, idb = P[sel], ids[sel]
        dpb, dib, dob = G.DeviceArray.from_host(Pb), G.DeviceArray.from_host(idb), G.DeviceArray(len(sel) * 8)
        run = lambda: G.check(lib.gwh_loglike_multi_device(idb.ctypes.data_as(ip), dib.ptr, G.IMRPHENOMD, G.MIXED,
                                                           dpb.ptr, len(sel), dob.ptr, None))
        run()
        G.synchronize()
        a = time.perf_counter()
        reps = 5
        for _ in range(reps):
            run()
            G.synchronize()
        el = (time.perf_counter() - a) / reps
        buckets[int(b)] = {"instances": int(len(sel)), "bins_per_instance_mean": float(nbins[sel].mean()),
                           "bins_per_lane": tiles[int(b)][0], "lanes": tiles[int(b)][1], "ms_per_launch": el * 1e3,
                           "bin_evals_per_s": float(nbins[sel].sum() / el)}
        for d in (dpb, dib, dob):
            d.free()
    rec = {"instances_per_launch": n, "events": len(events), "bins_min": int(nbins.min()),
           "bins_max": int(nbins.max()), "bins_total_per_launch": int(nbins.sum()),
           "resident_event_bytes": int(resident), "setup_s": t_setup, "autotune_s": t_tune,
           "autotuned_tiling": {int(i): {"bins_per_lane": b, "lanes": l} for i, (b, l) in tiles.items()},
           "mixed_evals_per_s": steps * n / wall, "mixed_bin_evals_per_s": steps * float(nbins.sum()) / wall,
           "ms_per_launch": wall / steps * 1e3, "p50_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
           "p99_ms": float(np.percentile(np.array(lat) * 1e3, 99)), "fp64_evals_per_s": fp64_rate,
           "max_abs_dlogL_mixed_vs_fp64": dmax, "per_bucket": buckets}
    log("uneven batch: %d instances, %d..%d bins, %.3f ms/launch, %.3e evals/s, %.3e bin-evals/s, max|dlogL| %.2e"
        % (n, nbins.min(), nbins.max(), wall / steps * 1e3, steps * n / wall, steps * nbins.sum() / wall, dmax))
    for d in (d_params, d_ids, d_mixed, d_fp64):
        d.free()
    for ev in events:
        ev.close()
    G.set_bins_per_lane(0)
    G.set_tile_lanes(0)
    return rec


def bench_fisher(G, W, ev, n_points=2048):
    """IMRPhenomPv2 13-parameter Fisher matrices at n_points precessing parameter points on the resident
    3-detector event: order-2 central differences with FP64 numerators, Gram on the FP64 matrix cores (the
    GWH_FP64 and GWH_MIXED modes), or FP32 operands on the FP32 matrix cores with FP64 accumulation across tiles
    (GWH_FISHER_GRAM_FP32), FP64 LU inversion -> marginal sigmas."""
    P = W.precessing_params(n_points, seed_stream=7)
    g = W.rng(3, 0, stream=8)
    P["chip"] = g.uniform(0.05, 0.9, n_points)
    P["phip"] = g.uniform(0, 2 * np.pi, n_points)
    for prec in (G.FP64, G.FISHER_GRAM_FP32):  # warm-up: each mode's buffers reach their size outside the timing
        G.fisher_batch(ev, P, G.IMRPHENOMPV2, G.BASIS_PV2_FULL, 13, precision=prec, with_sigma=True)
    out = {}
    res = {}
    for name, prec in (("fp64", G.FP64), ("fp32_gram", G.FISHER_GRAM_FP32)):
        a = time.perf_counter()
        F, sig = G.fisher_batch(ev, P, G.IMRPHENOMPV2, G.BASIS_PV2_FULL, 13, precision=prec, with_sigma=True)
        el = time.perf_counter() - a
        res[name] = (F, sig)
        out[name + "_points_per_s"] = n_points / el
        out[name + "_s"] = el
    (Fd, sd), (Fs, ss) = res["fp64"], res["fp32_gram"]
    # the mixed precision flag runs the FP64 Gram: its matrices are those of the FP64 mode
    Fm, sm = G.fisher_batch(ev, P[:64], G.IMRPHENOMPV2, G.BASIS_PV2_FULL, 13, precision=G.MIXED, with_sigma=True)
    dg = np.sqrt(np.abs(np.einsum("nii->ni", Fd)))
    norm = np.abs(Fs - Fd) / (dg[:, :, None] * dg[:, None, :])
    fin = np.isfinite(ss) & np.isfinite(sd)
    srel = np.abs(ss[fin] / sd[fin] - 1)
    fin64 = np.isfinite(sm) & np.isfinite(sd[:64])
    out.update({"points": n_points, "dim": 13, "bins": int(ev.n), "detectors": int(ev.ndet),
                "mixed_equals_fp64": bool(np.array_equal(Fm, Fd[:64])),
                "mixed_vs_fp64_sigma_rel_error_max": float(np.max(np.abs(sm[fin64] / sd[:64][fin64] - 1)))
                if fin64.any() else None,
                "fp32_gram_vs_fp64_max_normalised_element_error": float(norm.max()),
                "fp32_gram_vs_fp64_sigma_rel_error_median": float(np.median(srel)) if srel.size else None,
                "fp32_gram_vs_fp64_sigma_rel_error_max": float(srel.max()) if srel.size else None,
                "finite_sigma_fraction_fp64": float(np.isfinite(sd).mean())})
    log("Fisher (Pv2, 13 params, %d points): fp64 Gram %.1f points/s, fp32 Gram %.1f points/s (max normalised "
        "element error %.2e, sigma error median %.2e); mixed == fp64: %s"
        % (n_points, out["fp64_points_per_s"], out["fp32_gram_points_per_s"], norm.max(),
           out["fp32_gram_vs_fp64_sigma_rel_error_median"] or float("nan"), out["mixed_equals_fp64"]))
    return out


def bench_c0(G, W, O=None, cpu_threads=0, cpu_seconds=3.0):
    """The single-detector workload (Hanford, 4017 bins 20-1024 Hz at 1/4 Hz, 1024 parameter vectors around the
    injection) through the batched API, and its 7x7 sky-averaged IMRPhenomD Fisher at the injection as a single
    call: the reference's absolute step 1e-8 and the 1e-7 relative step.  With O (the oracle) the same work is
    timed on the host cores."""
    c = W.config0()
    f, psd = c["f"], c["psd"]
    inj = c["injection"]
    pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item()) for k in inj.dtype.names})
    data = [G.fourier_detector_response(f, pinj, 0) + W.noise(f, psd, c["noise_gen"])]
    ev = G.Event(f, data, [psd], [0])
    P = c["params"]
    out = {"instances": len(P), "bins": len(f)}
    for name, prec in (("fp64", G.FP64), ("mixed", G.MIXED)):
        G.loglike_batch(ev, P, G.IMRPHENOMD, prec)
        reps = 20
        a = time.perf_counter()
        for _ in range(reps):
            ll = G.loglike_batch(ev, P, G.IMRPHENOMD, prec)
        out[name + "_batched_evals_per_s"] = reps * len(P) / (time.perf_counter() - a)
    sa = dict((k, (list(inj[k]) if inj[k].shape else inj[k].item())) for k in inj.dtype.names)
    sa["sky_average"] = 1
    gp = G.make_params(**sa)
    for name, eps in (("abs_1e-8", 1e-8), ("rel_1e-7", -1e-7)):
        G.fisher_numerical(f, psd, G.BASIS_PHENOMD_SKYAVG, 0, 0, 7, gp, epsilon=eps)
        reps = 50
        a = time.perf_counter()
        for _ in range(reps):
            F = G.fisher_numerical(f, psd, G.BASIS_PHENOMD_SKYAVG, 0, 0, 7, gp, epsilon=eps)
        out["fisher_sa7_" + name + "_calls_per_s"] = reps / (time.perf_counter() - a)
        out["fisher_sa7_" + name + "_sigma"] = [float(x) for x in np.sqrt(np.abs(np.diag(np.linalg.inv(F))))]
    if O is not None:
        op = O.make_params(**sa)
        plist = [O.make_params(**{k: (list(P[i][k]) if P[i][k].shape else P[i][k].item()) for k in P.dtype.names})
                 for i in range(len(P))]
        a = time.perf_counter()
        m = 0
        while time.perf_counter() - a < cpu_seconds:
            O.network_loglike_batch(f, data, [psd], [0], plist, nthreads=cpu_threads)
            m += len(plist)
        out["cpu_evals_per_s"] = m / (time.perf_counter() - a)
        a = time.perf_counter()
        k = 0
        while time.perf_counter() - a < cpu_seconds / 3 or k < 2:
            O.fisher_numerical(f, psd, "IMRPhenomD", 0, 0, 7, op, epsilon=1e-8)
            k += 1
        out["cpu_fisher_sa7_calls_per_s_1thread"] = k / (time.perf_counter() - a)
    ev.close()
    log("c0: batched fp64 %.3e, mixed %.3e evals/s; SA Fisher %.0f calls/s (abs step), %.0f (relative step)%s"
        % (out["fp64_batched_evals_per_s"], out["mixed_batched_evals_per_s"], out["fisher_sa7_abs_1e-8_calls_per_s"],
           out["fisher_sa7_rel_1e-7_calls_per_s"],
           "; cpu %.3e evals/s, %.1f Fisher/s (1 thread)" % (out["cpu_evals_per_s"],
                                                            out["cpu_fisher_sa7_calls_per_s_1thread"])
           if O is not None else ""))
    return out


def bench_threaded_calls(G, W, f, data, psd, dets, inj, batched_rate, threads=(1, 32, 64), count=8192):
    """The reference-signature C++ layer called from host threads of one process (the threaded single-call
    driver, host/gwh_threaded_calls.cpp): FP64 MCMC_likelihood_extrinsic calls on the c1 arrays, coalesced by
    the engine into shared launches.  Calls per second per thread count and cache mode; every thread count's
    results must equal the one-thread results bit for bit."""
    import subprocess
    import tempfile
    exe = G.THREADED_CALLS
    if not os.path.exists(exe):
        raise RuntimeError("threaded-call driver not built (python -m gw_analysis_tools_amd.build)")
    P = W.perturbed_params(inj, count, step=7, scale=0.3)
    out = {"calls": count, "precision": "fp64", "runs": []}
    with tempfile.TemporaryDirectory() as tmp:
        inp = os.path.join(tmp, "in.bin")
        G.write_threaded_calls_input(inp, f, data, [psd] * len(dets), dets, P)
        ref = None
        for mode in ("pointer", "content"):
            for t in threads:
                if mode == "content" and t not in (1, max(threads)):
                    continue
                o = os.path.join(tmp, "out_%s_%d.bin" % (mode, t))
                r = subprocess.run([exe, inp, o, str(t), mode, "IMRPhenomD"], capture_output=True, text=True,
                                   timeout=300)
                if r.returncode != 0:
                    raise RuntimeError("driver failed: " + r.stderr.strip())
                rec = json.loads(r.stdout)
                res = np.fromfile(o, dtype=np.float64)
                if ref is None:
                    ref = res
                rec["bitwise_equal_to_one_thread"] = bool(np.array_equal(res, ref))
                rec["fraction_of_batched_fp64_evals_per_s"] = rec["calls_per_s"] / batched_rate
                out["runs"].append(rec)
                log("threaded single calls (%s keys, %d threads): %.3e calls/s (%.0f%% of batched FP64), %d shared "
                    "launches, up to %d calls per launch, equal to one thread: %s"
                    % (mode, t, rec["calls_per_s"], 100 * rec["fraction_of_batched_fp64_evals_per_s"],
                       rec["batches"], rec["max_calls_per_batch"], rec["bitwise_equal_to_one_thread"]))
    return out


def bench_pt_device(G, W, ev, inj, temps, n_chains, d_params, n, steps=200):
    """The whole parallel-tempering step on the device (proposal, prior, likelihood of every chain,
    Metropolis-Hastings, ladder swaps every 10th step) replayed from one HIP graph per step, against the
    host-driven step (numpy proposals and decisions, one engine launch per step) and the captured
    likelihood step alone (graph replay vs eager enqueue, each followed by a synchronisation)."""
    from gw_analysis_tools_amd import pt
    th0 = pt.theta_from_params(inj)
    widths = np.array([0.02, 0.02, 0.05, 0.03, 0.05, 2e-4, 0.03, 2e-4, 2e-3, 0.05, 0.05])
    out = {"instances": len(temps) * n_chains, "steps": steps, "swap_every": 10}
    for name, prec in (("fp64", G.FP64), ("mixed_guard", G.MIXED)):
        d = G.DevicePT(ev, inj, th0, widths, temps, n_chains, pt.LOWER, pt.UPPER, precision=prec, seed=1,
                       swap_every=10, steps_per_graph=10)
        d.run(20)
        a = time.perf_counter()
        d.run(steps)
        el = time.perf_counter() - a
        acc, sw, guard = d.counters()
        out["device_step_us_" + name] = el / steps * 1e6
        out["device_evals_per_s_" + name] = steps * d.n / el
        out["guard_fp64_evals_" + name] = int(guard)
        out["acceptance_" + name] = float(acc.sum()) / ((steps + 20) * d.n)
        d.close()
    for name, prec in (("fp64", G.FP64), ("mixed_guard", G.MIXED)):
        e = pt.PTEnsemble(ev, inj, th0, widths, temps, n_chains, precision=prec)
        e.run(3)
        hs = max(20, steps // 4)
        a = time.perf_counter()
        e.run(hs)
        out["host_step_us_" + name] = (time.perf_counter() - a) / hs * 1e6
    d_ll = G.DeviceArray(n * 8)
    lib = G.lib()
    for name, prec in (("fp64", G.FP64), ("mixed", G.MIXED)):
        g = G.LoglikeGraph(ev, d_params.ptr, n, d_ll.ptr, precision=prec)
        for _ in range(10):
            g.launch()
        g.synchronize()
        a = time.perf_counter()
        for _ in range(steps):
            g.launch()
            g.synchronize()
        out["graph_likelihood_step_us_" + name] = (time.perf_counter() - a) / steps * 1e6
        a = time.perf_counter()
        for _ in range(steps):
            G.check(lib.gwh_loglike_batch_device(None, ev.id, G.IMRPHENOMD, prec, d_params.ptr, n, d_ll.ptr, None))
            G.synchronize()
        out["eager_likelihood_step_us_" + name] = (time.perf_counter() - a) / steps * 1e6
        g.close()
    d_ll.free()
    log("device PT step: fp64 %.1f us, mixed+guard %.1f us (%d guard evals); host-driven %.1f / %.1f us; "
        "likelihood step graph %.1f / eager %.1f us (mixed)"
        % (out["device_step_us_fp64"], out["device_step_us_mixed_guard"], out["guard_fp64_evals_mixed_guard"],
           out["host_step_us_fp64"], out["host_step_us_mixed_guard"], out["graph_likelihood_step_us_mixed"],
           out["eager_likelihood_step_us_mixed"]))
    return out


def bench_maxll(G, W, ev, inj, n=512, reps=10):
    """The intrinsic MCMC likelihood (maximised over time, phase and distance with rocFFT transforms of length
    N per detector) of n IMRPhenomD and n IMRPhenomPv2 parameter sets on the c1 event: evaluations per second."""
    out = {"instances": n, "reps": reps}
    for name, model, P in (("imrphenomd", G.IMRPHENOMD, W.perturbed_params(inj, n, step=3)),
                           ("imrphenompv2", G.IMRPHENOMPV2, W.precessing_params(n))):
        G.maximized_loglike_batch(ev, P, model)
        a = time.perf_counter()
        for _ in range(reps):
            G.maximized_loglike_batch(ev, P, model)
        out[name + "_evals_per_s"] = reps * n / (time.perf_counter() - a)
    log("maximised (intrinsic) likelihood: IMRPhenomD %.3e, IMRPhenomPv2 %.3e evals/s"
        % (out["imrphenomd_evals_per_s"], out["imrphenompv2_evals_per_s"]))
    return out


def bench_serving(G, W, rates, seconds, fisher_fraction=0.1, clients=32, min_requests=1000):
    """Open-loop Poisson arrivals from `clients` threads at each aggregate rate, for max(seconds,
    min_requests / rate) seconds: a request is a log-likelihood batch of U{16..1024} instances (IMRPhenomD or
    IMRPhenomPv2, 50/50, mixed precision) on one of 64 resident events with the uneven-batch grid lengths, or
    (fisher_fraction) a 1..64-point FP64 Fisher batch of either model (IMRPhenomD 11-parameter or IMRPhenomPv2
    15-parameter MCMC basis); admission by a 200 GB HBM budget.  Latency = submission to completion per
    request; served rate = completed requests / (last completion - first arrival)."""
    import threading
    specs = W.config2(n_events=64, chains=4)
    g = W.rng(4, 1)
    events = []
    for sp in specs:
        f, psd = sp["f"], sp["psd"]
        inj = sp["injection"]
        p = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item()) for k in inj.dtype.names})
        data = [r + W.noise(f, psd, g) for r in G.create_coherent_GW_detection(f, p, sp["dets"])]
        events.append((G.Event(f, data, [psd] * 3, sp["dets"]), sp))
    rng = np.random.default_rng(20261204)
    pool = []  # pre-generated requests (parameter generation stays out of the arrival loop)
    for k in range(400):
        e, sp = events[rng.integers(len(events))]
        pv2 = rng.random() < 0.5
        if rng.random() < fisher_fraction:
            npts = int(rng.integers(1, 65))
            P = W.perturbed_params(sp["injection"], npts, step=k, config=4, scale=0.1)
            if pv2:
                P["spin1"][:, 0] = 0.2
                P["spin2"][:, 1] = -0.1
                pool.append(("fisher", e, P, G.IMRPHENOMPV2, (G.BASIS_PV2_MCMC, 15)))
            else:
                pool.append(("fisher", e, P, G.IMRPHENOMD, (G.BASIS_PHENOMD_MCMC, 11)))
        else:
            n = int(rng.integers(16, 1025))
            P = W.perturbed_params(sp["injection"], n, step=k, config=4, scale=0.1)
            if pv2:
                P["spin1"][:, 0] = 0.2
                P["spin2"][:, 1] = -0.1
            pool.append(("loglike", e, P, G.IMRPHENOMPV2 if pv2 else G.IMRPHENOMD, None))
    out = {"clients": clients, "events": len(events), "fisher_fraction": fisher_fraction,
           "hbm_budget_bytes": 200 << 30, "resident_event_bytes": int(sum(e.nbytes for e, _ in events)), "rates": []}
    br = G.Broker(hbm_budget_bytes=200 << 30, max_batch_instances=8192)
    # warm-up: every event and both models once
    for e, sp in events:
        for m in (G.IMRPHENOMD, G.IMRPHENOMPV2):
            t, _ = br.submit_loglike(e, W.perturbed_params(sp["injection"], 16, step=1, config=4, scale=0.1), m, G.MIXED)
            br.wait(t)
    for rate in rates:
        lock = threading.Lock()
        lats, done_t, kinds = [], [], []
        span = max(seconds, min_requests / rate)
        t_start = time.perf_counter() + 0.05

        def client(cid):
            r2 = np.random.default_rng(1000 * rate + cid)
            t = t_start
            mine = []
            while True:
                t += r2.exponential(clients / rate)
                if t - t_start > span:
                    break
                kind, e, P, model, fb = pool[r2.integers(len(pool))]
                dt = t - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                if kind == "fisher":
                    tk, _, _ = br.submit_fisher(e, P, model, fb[0], fb[1], precision=G.FP64)
                else:
                    tk, _ = br.submit_loglike(e, P, model, G.MIXED)
                mine.append((tk, kind))
            for tk, kind in mine:
                lat = br.wait(tk)
                with lock:
                    lats.append(lat)
                    kinds.append(kind)
                    done_t.append(time.perf_counter())

        threads = [threading.Thread(target=client, args=(c,)) for c in range(clients)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        wall = max(done_t) - t_start if done_t else float("nan")
        lat_ms = np.array(lats) / 1e3
        ll = lat_ms[np.array(kinds) == "loglike"]
        fi = lat_ms[np.array(kinds) == "fisher"]
        rec = {"rate_req_per_s": rate, "seconds": span, "requests": len(lats),
               "offered_req_per_s": len(lats) / span, "throughput_req_per_s": len(lats) / wall,
               "fisher_requests": int(fi.size),
               "fisher_p50_ms": float(np.percentile(fi, 50)) if fi.size else None,
               "fisher_p99_ms": float(np.percentile(fi, 99)) if fi.size else None,
               "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
               "loglike_p50_ms": float(np.percentile(ll, 50)) if ll.size else None,
               "loglike_p99_ms": float(np.percentile(ll, 99)) if ll.size else None}
        out["rates"].append(rec)
        log("serving %5d req/s offered (%.1f realised): %7.1f req/s served, p50 %.2f ms, p99 %.2f ms (%d requests, "
            "%d Fisher, Fisher p99 %.1f ms)"
            % (rate, rec["offered_req_per_s"], rec["throughput_req_per_s"], rec["p50_ms"], rec["p99_ms"], len(lats),
               fi.size, rec["fisher_p99_ms"] or float("nan")))
    out["broker"] = br.stats()
    # mixed vs FP64 on a sample of the likelihood requests
    worst = 0.0
    for kind, e, P, model, _ in pool[:60]:
        if kind != "loglike":
            continue
        t1, a = br.submit_loglike(e, P, model, G.MIXED)
        t2, b = br.submit_loglike(e, P, model, G.FP64)
        br.wait(t1)
        br.wait(t2)
        worst = max(worst, float(np.max(np.abs(a - b))))
    out["max_abs_dlogL_mixed_vs_fp64"] = worst
    br.close()
    for e, _ in events:
        e.close()
    return out


def cpu_threads():
    """host threads for the CPU legs: the CPUs this process may run on, at most 16 (a one-GPU box's share of its
    host cores)"""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(avail, 16, env if env > 0 else 16))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines(G, W, f, data, psd, dets, P, n, seconds):
    """The oracle (the reference algorithm restated in C, one instance per call) on the host cores: the headline
    workload (bounded sample, OpenMP over instances), a warm single-thread rate over at least one second,
    IMRPhenomPv2 on the same grid, and 13-parameter IMRPhenomPv2 network Fisher matrices (threads over points)."""
    import concurrent.futures
    import oracle as O
    cores = cpu_threads()

    def oparams(rec):
        return O.make_params(**{k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names})

    plist = [oparams(P[i]) for i in range(n)]
    O.network_loglike_batch(f, data, [psd] * 3, dets, plist[:2], nthreads=1)  # warm
    a = time.perf_counter()
    k = 0
    while time.perf_counter() - a < 1.0:
        O.network_loglike_batch(f, data, [psd] * 3, dets, plist[k % n:k % n + 1], nthreads=1)
        k += 1
    per_eval_1t = (time.perf_counter() - a) / k
    m = int(max(cores, seconds * cores / per_eval_1t))
    sample = [plist[i % n] for i in range(m)]
    a = time.perf_counter()
    _, used = O.network_loglike_batch(f, data, [psd] * 3, dets, sample, nthreads=cores)
    el = time.perf_counter() - a
    out = {"value": m / el, "unit": "logL evals/s", "cores": int(used), "kind": "port", "cpu_model": cpu_model(),
           "host_cpus_visible": os.cpu_count(),
           "sample": "%d evals of the headline workload (the first %d instances of the first batch, cycled), "
                     "oracle network_loglike with OpenMP over instances on %d threads (a one-GPU box's share)"
                     % (m, min(m, n), used),
           "seconds": el, "single_thread_ms_per_eval": per_eval_1t * 1e3, "single_thread_evals_timed": k}
    # IMRPhenomPv2 likelihood on the same grid
    P3 = W.precessing_params(64)
    p3 = [oparams(P3[i]) for i in range(len(P3))]
    O.network_loglike_batch(f, data, [psd] * 3, dets, p3[:1], model=O.IMRPHENOMPV2, nthreads=1)
    m3 = max(cores, int(seconds / 3 * cores / (3 * per_eval_1t)))
    a = time.perf_counter()
    O.network_loglike_batch(f, data, [psd] * 3, dets, [p3[i % 64] for i in range(m3)], model=O.IMRPHENOMPV2,
                            nthreads=cores)
    out["imrphenompv2_evals_per_s"] = m3 / (time.perf_counter() - a)
    # 13-parameter IMRPhenomPv2 network Fisher (the GPU Fisher leg's basis and points), threads over points
    PF = W.precessing_params(2048, seed_stream=7)
    g = W.rng(3, 0, stream=8)
    PF["chip"] = g.uniform(0.05, 0.9, 2048)
    PF["phip"] = g.uniform(0, 2 * np.pi, 2048)

    def fisher(i):
        op = oparams(PF[i])
        return sum(O.fisher_numerical(f, psd, "IMRPhenomPv2", d, dets[0], 13, op) for d in dets)

    fisher(0)
    npts = cores
    a = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(fisher, range(1, 1 + npts)))
    out["imrphenompv2_fisher13_points_per_s"] = npts / (time.perf_counter() - a)
    log("cpu baseline (%s, %d threads): %.1f evals/s (single thread %.2f ms/eval over %d evals); Pv2 %.1f evals/s; "
        "Pv2 13-parameter Fisher %.2f points/s"
        % (out["cpu_model"], used, out["value"], per_eval_1t * 1e3, k, out["imrphenompv2_evals_per_s"],
           out["imrphenompv2_fisher13_points_per_s"]))
    return out


KT_EVERY = 8  # per-bin kernel timed with HIP events on every KT_EVERY-th step of the timed region


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--bpl", type=int, default=0, help="bins per lane (0: autotuned)")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per tile (0: autotuned)")
    ap.add_argument("--legs", choices=["all", "mixed"], default="all",
                    help="mixed: only the timed mixed-precision leg (for counter runs)")
    ap.add_argument("--no-c2", action="store_true", help="skip the uneven multi-event batch leg")
    ap.add_argument("--no-fisher", action="store_true", help="skip the batched Fisher leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (broker) leg")
    ap.add_argument("--no-pt", action="store_true", help="skip the device parallel-tempering leg")
    ap.add_argument("--no-threads", action="store_true", help="skip the threaded single-call leg")
    ap.add_argument("--serving-rates", default="50,100,200,500,1000,2000")
    ap.add_argument("--serving-seconds", type=float, default=2.0)
    ap.add_argument("--c2-steps", type=int, default=20)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this size (reference point for traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("this engine measures one GPU: every GPU box here has one GPU; multi-GPU sharding is a design "
              "note only (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import gw_analysis_tools_amd as G
    from gw_analysis_tools_amd import workloads as W

    G.check(G.lib().gwh_init(0))
    if args.bpl:
        G.set_bins_per_lane(args.bpl)
    if args.lanes:
        G.set_tile_lanes(args.lanes)
    lib = G.lib()
    cfg = W.config1()
    f, psd, dets = cfg["f"], cfg["psd"], cfg["dets"]
    inj = W.injection_params(1)[0]
    pinj = G.make_params(**{k: (list(inj[k]) if inj[k].shape else inj[k].item())
                            for k in inj.dtype.names})
    resp = G.create_coherent_GW_detection(f, pinj, dets)
    g = W.rng(1, 0)
    data = [r + W.noise(f, psd, g) for r in resp]
    ev = G.Event(f, data, [psd] * 3, dets)
    n = len(cfg["temps"]) * cfg["n_chains"]
    K, Wm = args.steps, args.warmup
    total = K + Wm
    P = np.concatenate([W.perturbed_params(inj, n, step=s) for s in range(total)])
    psz = P.dtype.itemsize
    d_params = G.DeviceArray.from_host(P)
    d_mixed = G.DeviceArray(total * n * 8)
    d_fp64 = G.DeviceArray(total * n * 8)

    def enqueue(s, prec, out, model=G.IMRPHENOMD, dev_params=d_params, nn=n):
        G.check(lib.gwh_loglike_batch_device(None, ev.id, model, prec, dev_params.ptr + s * nn * psz, nn,
                                             out.ptr + s * nn * 8, None))

    # tiling of the c1 event (lanes per tile x bins per lane) by the autotuner, before the warm-up
    t0 = time.perf_counter()
    tiling = {"autotuned": not (args.bpl or args.lanes)}
    if tiling["autotuned"]:
        ids1 = np.full(n, ev.id, dtype=np.int32)
        d_ids1 = G.DeviceArray.from_host(ids1)
        choice = G.autotune_tiles(ids1, d_ids1, G.IMRPHENOMD, G.MIXED, d_params, d_mixed, reps=5)
        d_ids1.free()
        (bpl_c1, lanes_c1), = choice.values()
    else:
        bpl_c1, lanes_c1 = args.bpl or 8, args.lanes or 256
    tiling.update({"bins_per_lane": bpl_c1, "lanes": lanes_c1, "autotune_s": time.perf_counter() - t0})
    log("c1 tiling: %d lanes x %d bins per lane per tile (%s)"
        % (lanes_c1, bpl_c1, "autotuned" if tiling["autotuned"] else "set"))

    for s in range(Wm):
        enqueue(s, G.MIXED, d_mixed)
        G.synchronize()
    # ---- timed region: K steps of the mixed engine; HIP events around the per-bin kernel of every 8th step
    # (the event markers would otherwise add their own latency to every step's wall clock) ----
    G.synchronize()
    G.check(lib.gwh_kernel_timing_start(K))
    G.check(lib.gwh_kernel_timing_every(KT_EVERY))
    lat = []
    t0 = time.perf_counter()
    for s in range(Wm, total):
        ts = time.perf_counter()
        enqueue(s, G.MIXED, d_mixed)
        G.synchronize()
        lat.append(time.perf_counter() - ts)
    G.synchronize()
    t1 = time.perf_counter()
    ktot = np.zeros(1)
    kcnt = ctypes.c_int(0)
    G.check(lib.gwh_kernel_timing_stop(ktot.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(kcnt)))
    kernel_ms = float(ktot[0]) / max(kcnt.value, 1)
    wall = t1 - t0
    value = K * n / wall
    lat_ms = np.array(lat) * 1e3
    log("mixed: %.3f ms/step  %.3e evals/s  p50 %.3f ms  p99 %.3f ms; per-bin kernel %.4f ms avg over %d launches "
        "(every %d-th timed step)"
        % (wall / K * 1e3, value, np.percentile(lat_ms, 50), np.percentile(lat_ms, 99), kernel_ms, kcnt.value,
           KT_EVERY))
    # algorithmic work of the timed batches (frozen FLOP-equivalent table, per-instance region counts)
    fe_tot = nz_tot = 0.0
    for s0 in range(Wm, total, 64):
        s1 = min(total, s0 + 64)
        fe_c, nz_c = fe_per_launch(G.work_counts(ev, P[s0 * n:s1 * n], G.IMRPHENOMD), len(dets))
        fe_tot += fe_c
        nz_tot += nz_c
    fe = fe_tot / K
    nz = nz_tot / K
    achieved = fe / (kernel_ms * 1e-3) / 1e12
    bytes_alg = nz * (8 + 24 * len(dets))
    log("%.3e fe/launch -> %.2f TFLOP/s-eq (%.1f%% of %.1f); algorithmic bytes %.3e -> %.1f GB/s"
        % (fe, achieved, 100 * achieved / PEAK_VALU_TFLOPS, PEAK_VALU_TFLOPS, bytes_alg,
           bytes_alg / (kernel_ms * 1e-3) / 1e9))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "loglike_mixed_c1_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_VALU_TFLOPS, "traffic": traffic,
                "kernel": "gwh_loglike_grid_kernel<IMRPhenomD, mixed, %d, %d>" % (bpl_c1, lanes_c1),
                "kernel_ms": kernel_ms,
                "fe_per_launch": fe, "hbm_algorithmic_GBs": bytes_alg / (kernel_ms * 1e-3) / 1e9,
                "hbm_peak_GBs": PEAK_HBM_GBS}
    if args.copy_calibration_mib:
        nb = args.copy_calibration_mib << 20
        src, dst = G.DeviceArray(nb), G.DeviceArray(nb)
        for _ in range(3):
            G.check(lib.gwh_device_copy(dst.ptr, src.ptr, nb))
        log("copy calibration: 3 x gwh_copy_kernel of %d bytes" % nb)
        src.free()
        dst.free()
    if args.legs == "mixed":
        print(json.dumps({"metric": METRIC, "value": value, "unit": "logL evals/s", "n_gpus": 1, "steps": K,
                          "warmup": Wm, "ms_per_step": wall / K * 1e3, "legs": "mixed only",
                          "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
                          "tiling": tiling, "roofline": roofline}))
        return
    # ---- FP64 batched on the same batches (untimed for value) ----
    tf0 = time.perf_counter()
    for s in range(Wm, total):
        enqueue(s, G.FP64, d_fp64)
        G.synchronize()
    tf1 = time.perf_counter()
    fp64_rate = K * n / (tf1 - tf0)
    mx = d_mixed.to_host(np.float64, total * n)[Wm * n:]
    fx = d_fp64.to_host(np.float64, total * n)[Wm * n:]
    dmax = float(np.max(np.abs(mx - fx)))
    log("fp64 batched: %.3e evals/s ; max |dlogL| mixed vs fp64 over %d evals: %.3e" % (fp64_rate, len(mx), dmax))
    # ---- plain FP64 one-instance-per-launch ----
    reps = min(n, 256)
    ts0 = time.perf_counter()
    for i in range(reps):
        G.check(lib.gwh_loglike_batch_device(None, ev.id, G.IMRPHENOMD, G.FP64, d_params.ptr + i * psz, 1,
                                             d_fp64.ptr + i * 8, None))
    G.synchronize()
    single_rate = reps / (time.perf_counter() - ts0)
    log("fp64 one-instance-per-launch: %.3e evals/s" % single_rate)
    # ---- launch breakdown (separate event-timed loop, not part of value) ----
    ms = np.zeros(1)
    kms = np.zeros(1)
    dp = ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    kp = kms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    G.check(lib.gwh_time_loglike_batch_device(ev.id, G.IMRPHENOMD, G.MIXED, d_params.ptr, n, d_mixed.ptr, 50, dp, kp))
    launch_ms = float(ms[0])
    G.check(lib.gwh_time_loglike_batch_device(ev.id, G.IMRPHENOMD, G.FP64, d_params.ptr, n, d_fp64.ptr, 20, dp, kp))
    kernel_ms_fp64 = float(kms[0])
    log("mixed launch (setup + per-bin kernel + finalize) %.4f ms; FP64 per-bin kernel %.4f ms"
        % (launch_ms, kernel_ms_fp64))
    roofline["launch_ms_setup_kernel_finalize"] = launch_ms
    roofline["kernel_ms_fp64"] = kernel_ms_fp64

    # ---- IMRPhenomPv2 on the same grid: 1024 chains, mixed batched vs plain FP64 single launches ----
    pv2 = {}
    try:
        n3 = 1024
        P3 = W.precessing_params(n3)
        d3 = G.DeviceArray.from_host(P3)
        o3 = G.DeviceArray(n3 * 8)
        o3b = G.DeviceArray(n3 * 8)
        for _ in range(3):
            enqueue(0, G.MIXED, o3, G.IMRPHENOMPV2, d3, n3)
            G.synchronize()
        reps3 = 20
        a = time.perf_counter()
        for _ in range(reps3):
            enqueue(0, G.MIXED, o3, G.IMRPHENOMPV2, d3, n3)
            G.synchronize()
        pv2_mixed = reps3 * n3 / (time.perf_counter() - a)
        enqueue(0, G.FP64, o3b, G.IMRPHENOMPV2, d3, n3)
        G.synchronize()
        pd = float(np.max(np.abs(o3.to_host(np.float64, n3) - o3b.to_host(np.float64, n3))))
        a = time.perf_counter()
        for _ in range(reps3):
            enqueue(0, G.FP64, o3b, G.IMRPHENOMPV2, d3, n3)
            G.synchronize()
        pv2_fp64 = reps3 * n3 / (time.perf_counter() - a)
        r1 = 256
        a = time.perf_counter()
        for i in range(r1):
            G.check(lib.gwh_loglike_batch_device(None, ev.id, G.IMRPHENOMPV2, G.FP64, d3.ptr + i * psz, 1,
                                                 o3b.ptr + i * 8, None))
        G.synchronize()
        pv2_single = r1 / (time.perf_counter() - a)
        pv2 = dict(mixed_batched_evals_per_s=pv2_mixed, fp64_batched_evals_per_s=pv2_fp64,
                   fp64_one_instance_per_launch_evals_per_s=pv2_single,
                   speedup_mixed_batched_vs_fp64_single=pv2_mixed / pv2_single, max_abs_dlogL_mixed_vs_fp64=pd)
        log("Pv2: mixed %.3e, fp64 batched %.3e, fp64 single %.3e evals/s, speedup %.1fx, max|dlogL| %.2e"
            % (pv2_mixed, pv2_fp64, pv2_single, pv2_mixed / pv2_single, pd))
    except Exception as e:  # reported, not fatal for the headline
        log("Pv2 leg failed:", e)
        pv2 = {"error": str(e)}

    # ---- batched Fisher matrices (IMRPhenomPv2, 13 parameters, 2048 points) on the same event ----
    fisher = None
    if not args.no_fisher:
        try:
            fisher = bench_fisher(G, W, ev)
        except Exception as e:  # reported, not fatal for the headline
            log("Fisher leg failed:", e)
            fisher = {"error": str(e)}

    # ---- the whole PT step on the device, one graph replay per step ----
    ptd = None
    if not args.no_pt:
        try:
            ptd = bench_pt_device(G, W, ev, inj, cfg["temps"], cfg["n_chains"], d_params, n)
        except Exception as e:  # reported, not fatal for the headline
            log("device PT leg failed:", e)
            ptd = {"error": str(e)}

    # ---- FFT-maximised (intrinsic) likelihood on the same event ----
    mxl = None
    if not args.no_pt:
        try:
            mxl = bench_maxll(G, W, ev, inj)
        except Exception as e:  # reported, not fatal for the headline
            log("maximised-likelihood leg failed:", e)
            mxl = {"error": str(e)}

    # ---- uneven batch: 256 events of 1e3..2.6e5 bins, 8 chains each, one launch per step ----
    c2 = None
    if not args.no_c2:
        try:
            c2 = bench_uneven(G, W, args.c2_steps)
        except Exception as e:  # reported, not fatal for the headline
            log("uneven-batch leg failed:", e)
            c2 = {"error": str(e)}

    # ---- serving: Poisson request streams from 32 clients through the continuous-batching broker ----
    serving = None
    if not args.no_serving:
        try:
            serving = bench_serving(G, W, [int(x) for x in args.serving_rates.split(",")], args.serving_seconds)
        except Exception as e:  # reported, not fatal for the headline
            log("serving leg failed:", e)
            serving = {"error": str(e)}

    # ---- the single-detector workload and its sky-averaged Fisher (GPU; CPU beside it unless --no-cpu) ----
    c0 = None
    try:
        O0 = None
        if not args.no_cpu:
            import oracle as O0
        c0 = bench_c0(G, W, O0, cpu_threads=cpu_threads())
    except Exception as e:  # reported, not fatal for the headline
        log("c0 leg failed:", e)
        c0 = {"error": str(e)}

    # ---- the reference-signature C++ layer called from many host threads (coalesced single calls) ----
    thr = None
    if not args.no_threads:
        try:
            thr = bench_threaded_calls(G, W, f, data, psd, dets, inj, fp64_rate)
        except Exception as e:  # reported, not fatal for the headline
            log("threaded-call leg failed:", e)
            thr = {"error": str(e)}

    # ---- CPU baseline: the oracle (reference algorithm, C, one instance per call) ----
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baselines(G, W, f, data, psd, dets, P, n, args.cpu_seconds)
        except Exception as e:
            log("cpu baseline failed:", e)
            cpu = {"error": str(e)}

    rec = {
        "metric": METRIC, "value": value, "unit": "logL evals/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": wall / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (mixed: f32 merger-ringdown exp)", "data": "synthetic",
        "config": {"workload": "PT ensemble step: IMRPhenomD, H1/L1/V1, 16225 bins 20-2048 Hz df=1/8 Hz, "
                               "16 temperatures x 32 chains = 512 instances per step",
                   "instances_per_step": n, "bins": len(f), "detectors": len(dets)},
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "fp64_batched_evals_per_s": fp64_rate, "fp64_one_instance_per_launch_evals_per_s": single_rate,
        "speedup_mixed_vs_fp64_one_instance_per_launch": value / single_rate,
        "max_abs_dlogL_mixed_vs_fp64": dmax,
        "tiling": tiling,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "imrphenompv2_c3_likelihood": pv2,
        "uneven_batch_c2": c2,
        "fisher_c3": fisher,
        "serving_c4": serving,
        "pt_step_on_device_c1": ptd,
        "maximized_intrinsic_likelihood_c1": mxl,
        "single_detector_c0": c0,
        "threaded_single_calls_c1": thr,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
