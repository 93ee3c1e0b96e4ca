"""Pins of the CPU restatement of the FFT-maximised likelihoods (oracle/gwo_maxll.py) against what the
mathematics fixes: the transform is the plain DFT, the statistics are invariant under the symmetries they
maximise over, they equal brute-force maximisations over the extrinsic parameters, and the coalescence
variant recovers a planted time shift."""
import numpy as np

import oracle as O
from oracle import gwo_maxll as M


def _toy(n=257, seed=1):
    g = np.random.default_rng(seed)
    f = 20. + 0.25 * np.arange(n)
    psd = 1. + g.random(n)
    hp = (g.normal(size=n) + 1j * g.normal(size=n)) * np.exp(-f / 60.)
    hc = (g.normal(size=n) + 1j * g.normal(size=n)) * np.exp(-f / 60.)
    d = g.normal(size=n) + 1j * g.normal(size=n)
    return f, psd, hp, hc, d


def test_transform_is_the_plain_dft():
    g = np.random.default_rng(3)
    x = g.normal(size=33) + 1j * g.normal(size=33)
    k = np.arange(33)
    direct = np.array([np.sum(x * np.exp(-2j * np.pi * k * m / 33)) for m in range(33)])
    assert np.allclose(np.fft.fft(x), direct, rtol=0, atol=1e-12)


def test_aligned_statistic_symmetries_and_closed_form():
    f, psd, hp, hc, d = _toy()
    n = len(f)
    base = M.aligned(d, psd, f, hp)
    # the statistic maximises over amplitude, phase and (circular) time of arrival
    assert np.isclose(M.aligned(d, psd, f, (0.3 - 2.1j) * hp), base, rtol=1e-12)
    shift = np.exp(-2j * np.pi * np.arange(n) * 37 / n)
    assert np.isclose(M.aligned(d * shift, psd, f, hp), base, rtol=1e-12)
    # noise-free data c h: the peak is at zero lag with value 0.5 |c|^2 (4 df sum |h|^2/S)^2 / (h|h)
    c = 2.5 - 0.5j
    df = f[1] - f[0]
    hh = 4 * O.simpsons_sum(df, np.abs(hp) ** 2 / psd)
    rect = 4 * df * np.sum(np.abs(hp) ** 2 / psd)
    assert np.isclose(M.aligned(c * hp, psd, f, hp), 0.5 * abs(c) ** 2 * rect ** 2 / hh, rtol=1e-12)
    # and equals the brute-force maximum of Re(2 A e^{i phi} rho_k) - A^2 (h|h) over A, phi, k
    rho = df * np.fft.fft(4 * np.conj(d) * hp / psd)
    A = np.linspace(0, 2 * np.abs(rho).max() / hh, 2001)[:, None]
    best = max(np.max(2 * A * np.abs(r) - A * A * hh) for r in rho)
    assert np.isclose(best / 2, base, rtol=1e-5)


def test_coal_statistic_recovers_a_planted_shift():
    f, psd, hp, hc, d = _toy()
    n = len(f)
    m = 41
    data = 3.0 * hp * np.exp(-2j * np.pi * np.arange(n) * m / n) * np.exp(0.7j)
    ll, tc, phi = M.coal(data, psd, f, hp)
    df = f[n // 2] - f[n // 2 - 1]
    assert np.isclose(tc, m / n / df) or np.isclose(tc, (n - m) / n / df)
    assert abs(np.angle(np.exp(1j * (phi + 0.7)))) < 1e-9 or abs(np.angle(np.exp(1j * (phi - 0.7)))) < 1e-9
    hh = 4 * O.simpsons_sum(df, np.abs(hp) ** 2 / psd)
    rect = 4 * df * np.sum(np.abs(hp) ** 2 / psd)
    assert np.isclose(ll, -0.5 * (hh - 2 * 3.0 * rect), rtol=1e-12)


def test_unaligned_statistic_is_the_maximum_over_polarisation():
    """0.25 lambda_k equals max over psi of |rho(psi)|^2 / (2 N(psi)), the log-likelihood ratio maximised over
    amplitude and phase of h = cos(psi) h+^ + sin(psi) hx^ (arXiv:1603.02444)"""
    f, psd, hp, hc, d = _toy(seed=4)
    df = f[1] - f[0]
    simp = lambda x: O.simpsons_sum(df, x)
    hpn = hp / np.sqrt(4 * simp(np.abs(hp) ** 2 / psd))
    hcn = hc / np.sqrt(4 * simp(np.abs(hc) ** 2 / psd))
    rp = df * np.fft.fft(4 * np.conj(d) * hpn / psd)
    rc = df * np.fft.fft(4 * np.conj(d) * hcn / psd)
    ipc = 4 * simp(np.real(hpn * np.conj(hcn)) / psd)
    psi = np.linspace(0, np.pi, 20001)[:, None]
    rho = np.cos(psi) * rp[None, :] + np.sin(psi) * rc[None, :]
    norm = 1 + 2 * np.sin(psi) * np.cos(psi) * ipc
    brute = np.max(np.abs(rho) ** 2 / (2 * norm))
    assert np.isclose(M.unaligned(d, psd, f, hp, hc), brute, rtol=1e-6)
    # a plus-only template reduces to the aligned statistic's maximum over amplitude and phase
    assert M.unaligned(d, psd, f, hp, hc) >= M.aligned(d, psd, f, hp) * (1 - 1e-12)
