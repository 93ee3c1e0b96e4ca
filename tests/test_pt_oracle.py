"""Pins of the CPU restatement of the device tempering step (oracle/pt_oracle.py): the Philox4x32-10
generator against the Random123 known-answer vectors, the deviates against their distributions, and the
Metropolis-Hastings / swap rules against the tempered target they must sample."""
import numpy as np
import pytest

from oracle import pt_oracle as PO


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key) -> output
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, out in kat:
        got = PO.philox4x32_10(*[np.array([c], dtype=np.uint64) for c in ctr], *key)
        assert [int(g[0]) for g in got] == list(out)


def test_deviates_have_their_distributions():
    n, dim = 200000, 11
    beta, z, logu = PO.mh_draws(20261203, 7, n, dim)
    assert beta.min() == 0 and beta.max() == dim - 1
    counts = np.bincount(beta, minlength=dim)
    assert np.all(np.abs(counts - n / dim) < 5 * np.sqrt(n / dim))
    assert abs(z.mean()) < 5 / np.sqrt(n) and abs(z.var() - 1) < 5 * np.sqrt(2 / n)
    assert abs(np.mean(z ** 4) - 3) < 0.1  # Gaussian kurtosis
    u = np.exp(logu)
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / n) and u.min() >= 0 and u.max() < 1
    a = PO.swap_draws(20261203, 7, 4, 1000)
    assert a.shape == (3, 1000) and abs(a.mean() - 0.5) < 0.02
    # streams differ by step, instance and tag
    b2, z2, _ = PO.mh_draws(20261203, 8, 16, dim)
    assert not np.array_equal(z[:16], z2)


def test_tempered_chains_sample_their_targets():
    """Gaussian likelihood with unit variance: the chains at temperature T sample N(0, T) when the MH and
    swap rules are right (a missing 1/T, a sign error or an always-accepted swap moves the variances)."""
    temps = np.array([1.0, 2.5])
    nc, dim = 96, 2

    def ll(th):
        return -0.5 * np.sum(np.atleast_2d(th) ** 2, axis=1)

    pt = PO.PTOracle(ll, np.zeros(dim), np.full(dim, 1.2), temps, nc, np.full(dim, -50.), np.full(dim, 50.),
                     seed=11, swap_every=2)
    samples = [[], []]
    swaps = 0
    for s in range(1500):
        pt.step()
        swaps += pt.history[-1]["swapped"].sum()
        pt.history.clear()
        if s >= 300:
            for k in range(2):
                samples[k].append(pt.theta[k * nc:(k + 1) * nc].copy())
    for k, T in enumerate(temps):
        x = np.concatenate(samples[k])
        assert abs(x.mean()) < 0.08, (k, x.mean())
        assert abs(x.var() / T - 1) < 0.08, (k, x.var(), T)
    assert swaps > 0


def test_out_of_prior_and_nan_reject():
    def ll(th):
        th = np.atleast_2d(th)
        return np.where(th[:, 0] > 0.5, np.nan, 0.0)

    pt = PO.PTOracle(ll, np.full(1, 0.45), np.full(1, 0.2), [1.0], 64, [0.0], [1.0], seed=3, swap_every=0)
    for _ in range(50):
        pt.step()
        h = pt.history[-1]
        assert not np.any(h["accept"] & ((h["prop"][:, 0] > 0.5) | (h["prop"][:, 0] < 0)))
    assert np.all((pt.theta >= 0) & (pt.theta <= 0.5))
