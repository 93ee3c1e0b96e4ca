"""The serving broker on the GPU: concurrent client threads submit likelihood requests of mixed models,
precisions, sizes and events plus Fisher requests; every result equals the same request run alone
(bit for bit), requests are merged into shared batches, and the HBM admission budget is honoured."""
import threading

import numpy as np
import pytest

import gw_analysis_tools_amd as G
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu


def _events(k=6):
    specs = W.config2(n_events=24, chains=4)[:k]
    g = W.rng(4, 0)
    out = []
    for sp in specs:
        f, psd = sp["f"], sp["psd"]
        inj = sp["injection"]
        p = G.make_params(**{n: (list(inj[n]) if inj[n].shape else inj[n].item())
                             for n in inj.dtype.names})
        data = [r + W.noise(f, psd, g) for r in G.create_coherent_GW_detection(f, p, sp["dets"])]
        out.append((G.Event(f, data, [psd] * 3, sp["dets"]), sp))
    return out


def _request(rng, evs):
    k = rng.integers(len(evs))
    n = int(rng.integers(8, 200))
    model = G.IMRPHENOMD if rng.random() < 0.5 else G.IMRPHENOMPV2
    prec = G.FP64 if rng.random() < 0.5 else G.MIXED
    base = evs[k][1]["injection"]
    P = W.perturbed_params(base, n, step=int(rng.integers(1 << 30)), config=4, scale=0.2)
    if model == G.IMRPHENOMPV2:
        P["spin1"][:, 0] = 0.2
        P["spin2"][:, 1] = -0.1
    if rng.random() < 0.3:  # a request spanning two events
        ids = [evs[k][0].id] * (n // 2) + [evs[(k + 1) % len(evs)][0].id] * (n - n // 2)
    else:
        ids = [evs[k][0].id] * n
    return ids, P, model, prec


def test_concurrent_requests_equal_alone():
    evs = _events()
    br = G.Broker(hbm_budget_bytes=200 << 30, max_batch_instances=4096)
    results = []
    lock = threading.Lock()

    def client(seed):
        rng = np.random.default_rng(seed)
        mine = []
        for _ in range(12):
            ids, P, model, prec = _request(rng, evs)
            t, out = br.submit_loglike(ids, P, model, prec)
            mine.append((t, out, ids, P, model, prec))
        for t, out, ids, P, model, prec in mine:
            lat = br.wait(t)
            with lock:
                results.append((out.copy(), ids, P, model, prec, lat))

    threads = [threading.Thread(target=client, args=(s,)) for s in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    st = br.stats()
    print("broker:", st, "latency p50 %.0f us p99 %.0f us" % (np.percentile([r[5] for r in results], 50),
                                                               np.percentile([r[5] for r in results], 99)))
    assert len(results) == 96
    assert st["max_requests_per_batch"] > 1  # requests were coalesced
    for out, ids, P, model, prec, _ in results:
        alone = G.loglike_batch(ids[0], P, model, prec, events=ids)
        assert np.array_equal(out, alone)
    br.close()
    for e, _ in evs:
        e.close()


def test_fisher_requests_and_admission_budget():
    evs = _events(3)
    e0, sp = evs[0]
    resident = sum(e.nbytes for e, _ in evs)
    # a budget that leaves room for small batches and a one-point Fisher pass only: big requests wait for
    # in-flight work, Fisher passes shrink, what cannot fit fails
    one_point = 3 * e0.n * 2 * 16 * 8 * 2  # FP64 operand rows of one point, with room for the Gram partials
    budget = resident + one_point + (8 << 20)
    br = G.Broker(hbm_budget_bytes=budget, max_batch_instances=4096)
    P = W.perturbed_params(sp["injection"], 4, step=3, config=4, scale=0.1)
    tf, F, _ = br.submit_fisher(e0, P, G.IMRPHENOMD, G.BASIS_PHENOMD_MCMC, 11)
    tickets = [br.submit_loglike(e0, W.perturbed_params(sp["injection"], 64, step=s, config=4, scale=0.1))
               for s in range(10)]
    br.wait(tf)
    ref = G.fisher_batch(e0, P, G.IMRPHENOMD, G.BASIS_PHENOMD_MCMC, 11)
    assert np.array_equal(F, ref)
    for t, out in tickets:
        br.wait(t)
        assert np.all(np.isfinite(out))
    st = br.stats()
    print("small-budget broker:", st)
    assert st["peak_device_bytes"] <= budget
    br.close()
    # a request that cannot fit even on an idle device is refused: 60000 instances on this 75k-bin event hold
    # ~5 kB of instance record and tile partials each, more than the budget leaves beside the events
    assert e0.n > 50000
    br = G.Broker(hbm_budget_bytes=budget, max_batch_instances=60000)
    huge = W.perturbed_params(sp["injection"], 60000, step=99, config=4, scale=0.1)
    t, _ = br.submit_loglike(e0, huge)
    with pytest.raises(G.GWHError):
        br.wait(t)
    br.close()
    for e, _ in evs:
        e.close()


def test_events_held_until_tickets_are_collected_and_fisher_merging():
    """an event a request reads cannot be destroyed until the ticket is waited for; concurrent Fisher requests of
    the same event, model and basis are merged into shared batches, each result equal to the request alone"""
    evs = _events(2)
    e0, sp = evs[0]
    br = G.Broker(hbm_budget_bytes=200 << 30, max_batch_instances=4096)
    reqs = []
    for s in range(8):
        P = W.perturbed_params(sp["injection"], 1 + s % 3, step=40 + s, config=4, scale=0.1)
        t, F, _ = br.submit_fisher(e0, P, G.IMRPHENOMD, G.BASIS_PHENOMD_MCMC, 11)
        reqs.append((t, F, P))
    tl, _ = br.submit_loglike(evs[1][0], W.perturbed_params(evs[1][1]["injection"], 16, step=1, config=4, scale=0.1))
    rc = G.lib().gwh_event_destroy(e0.id)
    assert rc == G.GWH_ERR_BUSY
    assert G.lib().gwh_event_destroy(evs[1][0].id) == G.GWH_ERR_BUSY
    for t, F, P in reqs:
        br.wait(t)
        assert np.array_equal(F, G.fisher_batch(e0, P, G.IMRPHENOMD, G.BASIS_PHENOMD_MCMC, 11))
    st = br.stats()
    print("Fisher merging:", st)
    assert st["fisher_requests"] == 8 and st["fisher_batches"] <= 8
    # collected: the Fisher requests' event is free again; the likelihood request still holds the other
    e0.close()
    assert G.lib().gwh_event_destroy(evs[1][0].id) == G.GWH_ERR_BUSY
    br.wait(tl)
    evs[1][0].close()
    br.close()
