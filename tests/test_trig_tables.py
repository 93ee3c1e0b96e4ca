"""The tables behind the kernels' table-driven sin/cos and exponential (gw_analysis_tools_amd/csrc/gwh_trig_tables.h)
and the two algorithms over them (gwh_sincos_table, gwh_exp_table in gwh_trig.h), restated here in numpy with the
same constants and operation order, against 50-digit mpmath values:

* every table entry is the correctly rounded double of cos / sin(2 pi j / 256) and 2^(j / 64);
* sin and cos through the table are within 4 ulp of 1 of the exact values of the reduced phase, over phases up
  to 1e6 rad (the per-bin phases of the likelihood reach ~1e5-1e6 rad);
* e^x through the table is within 1e-15 relative over [-700, 5];
* a planted error (one table entry off by ~1e-13, a wrong factorial in the series, one 2^(j/64) entry off by
  1e-12 relative) is caught;
* the arctangent seeded in FP32 and corrected once through the table (gwh_atan_seeded) is within 8 ulp of pi/2
  of the exact value for |x| up to 1e6, with any FP32 seed within 2^-22 relative of atan(x); a correction with
  the wrong sign is caught;
* the reciprocal refined by one Newton step (rcp_newton) squares its seed's relative error (within 2 ulp of 1/z
  for seeds within 2^-26, as the hardware FP64 reciprocal's), and the reciprocal square root from an FP32 seed
  and two Newton steps (rsqrt_seeded) is within 4 ulp of z^(-1/2) over [1, 1e12] for any seed within 2^-22; a
  step with the wrong sign, or one step short, is caught;
* the IMRPhenomPv2 half angle sin(b/2) = |s| cos b / (2 cos(b/2)) with cos b = (1 + s^2)^(-1/2) is within 4 ulp of
  sin(atan(s) / 2) for s from 1e-8 to 1e4, where sqrt((1 - cos b) / 2) loses digits as s -> 0."""
import os
import re

import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")
mpmath.mp.dps = 50

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gw_analysis_tools_amd", "csrc",
                   "gwh_trig_tables.h")


def _tables():
    text = open(HDR).read()
    turn = re.search(r"gwh_turn_table\[256\] = \{(.*?)\n\};", text, re.S).group(1)
    pairs = re.findall(r"\{([^,}]+), ([^}]+)\}", turn)
    tab = np.array([[float.fromhex(c.strip()), float.fromhex(s.strip())] for c, s in pairs])
    e2 = re.search(r"gwh_exp2_table\[64\] = \{(.*?)\n\};", text, re.S).group(1)
    exp2 = np.array([float.fromhex(v.strip()) for v in e2.split(",") if v.strip()])
    return tab, exp2


def _fma(a, b, c):
    """fused multiply-add of doubles: exact product and sum in 60-digit arithmetic, rounded once"""
    with mpmath.workprec(200):
        return float(mpmath.mpf(a) * mpmath.mpf(b) + mpmath.mpf(c))


def sincos_table(phase, tab, c3=-1.0 / 6.0):
    """gwh_sincos_table, operation by operation"""
    r = phase * 0.15915494309189533577
    k_round = 6755399441055744.0
    rb = _fma(r, 256.0, k_round)
    n = rb - k_round
    j = int(np.float64(rb).view(np.int64)) & 255
    y = _fma(n, -0.00390625, r)
    t = y * 6.283185307179586477
    t2 = t * t
    sl = t * _fma(t2, _fma(t2, 1.0 / 120.0, c3), 1.0)
    cl = _fma(t2, _fma(t2, _fma(t2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0)
    C, S = tab[j]
    return _fma(S, cl, C * sl), _fma(C, cl, -S * sl), r


def exp_table(x, exp2):
    """gwh_exp_table, operation by operation"""
    n = np.rint(x * 92.332482616893656877)
    r = _fma(-n, 0.01083042469326756, x)
    r = _fma(-n, 2.9815858269852933e-12, r)
    p = _fma(r, _fma(r, _fma(r, _fma(r, _fma(r, 1.0 / 120.0, 1.0 / 24.0), 1.0 / 6.0), 0.5), 1.0), 1.0)
    ni = int(n)
    return float(np.ldexp(exp2[ni & 63] * p, ni >> 6))


def test_entries_are_correctly_rounded():
    tab, exp2 = _tables()
    assert tab.shape == (256, 2) and exp2.shape == (64,)
    for j in range(256):
        a = 2 * mpmath.pi * j / 256
        assert tab[j, 0] == float(mpmath.cos(a)) and tab[j, 1] == float(mpmath.sin(a)), j
    for j in range(64):
        assert exp2[j] == float(mpmath.power(2, mpmath.mpf(j) / 64)), j


def _sincos_worst(tab, phases, c3=-1.0 / 6.0):
    worst = 0.0
    for x in phases:
        s, c, r = sincos_table(float(x), tab, c3)
        th = 2 * mpmath.pi * (mpmath.mpf(r) - mpmath.nint(mpmath.mpf(r)))  # the phase as reduced to turns
        worst = max(worst, abs(s - float(mpmath.sin(th))), abs(c - float(mpmath.cos(th))))
    return worst


def test_sincos_through_the_table():
    tab, _ = _tables()
    g = np.random.default_rng(7)
    phases = np.concatenate([g.uniform(-4, 4, 300), g.uniform(-1e6, 1e6, 300),
                             (np.arange(-64, 65) / 128.0) * np.pi])  # table nodes and midpoints
    worst = _sincos_worst(tab, phases)
    assert worst <= 4 * 2.0 ** -52, worst
    # planted errors
    bad = tab.copy()
    bad[37, 1] = np.nextafter(bad[37, 1], 2.0) + 1e-13
    assert _sincos_worst(bad, [2 * np.pi * 37 / 256 + 1e-3]) > 1e-14
    assert _sincos_worst(tab, phases[:50], c3=-1.0 / 5.0) > 1e-8  # a wrong factorial in the series


def test_exp_through_the_table():
    _, exp2 = _tables()
    g = np.random.default_rng(8)
    xs = np.concatenate([g.uniform(-700, 5, 400), g.uniform(-1, 1, 200), [0.0, -1e-300, np.log(2) / 128]])
    worst = 0.0
    for x in xs:
        ref = mpmath.exp(mpmath.mpf(float(x)))
        worst = max(worst, float(abs((mpmath.mpf(exp_table(float(x), exp2)) - ref) / ref)))
    assert worst <= 1e-15, worst
    bad = exp2.copy()
    bad[5] *= 1 + 1e-12
    x = 5 * np.log(2) / 64
    assert abs(exp_table(x, bad) / float(mpmath.exp(x)) - 1) > 1e-13


def atan_seeded(x, tab, sign=1.0):
    """gwh_atan_seeded, operation by operation (the FP32 seed: numpy's float32 arctangent)"""
    t0 = float(np.arctan(np.float32(x)))
    sn, cs, _ = sincos_table(t0, tab)
    num = _fma(x, cs, -sn)
    den = _fma(x, sn, cs)
    return t0 + sign * num * (1.0 / den)


def test_atan_seeded():
    tab, _ = _tables()
    g = np.random.default_rng(9)
    xs = np.concatenate([g.uniform(-2, 2, 300), g.uniform(-1e3, 1e3, 200), g.uniform(-1e6, 1e6, 100),
                         [0.0, 1e-300, 1.0, -1.0, 1e6, -1e6, 3.0e-8]])
    worst = max(abs(atan_seeded(float(x), tab) - float(mpmath.atan(mpmath.mpf(float(x))))) for x in xs)
    assert worst <= 8 * 2.0 ** -52 * np.pi / 2, worst
    # a seed perturbed within FP32 accuracy is corrected all the same
    x = 0.7
    t0 = float(np.arctan(np.float32(x))) * (1 + 2.0 ** -22)
    sn, cs, _ = sincos_table(t0, tab)
    fixed = t0 + _fma(x, cs, -sn) * (1.0 / _fma(x, sn, cs))
    assert abs(fixed - float(mpmath.atan(x))) < 1e-15
    # planted: the correction subtracted instead of added
    assert max(abs(atan_seeded(float(x), tab, -1.0) - float(mpmath.atan(mpmath.mpf(float(x))))) for x in xs[:50]) > 1e-9


def rcp_newton(z, seed_err):
    """rcp_newton, operation by operation, from a seed with relative error seed_err"""
    r = (1.0 / z) * (1.0 + seed_err)
    return _fma(r, _fma(-z, r, 1.0), r)


def rsqrt_seeded(z, seed_err, steps=2, sign=1.0):
    """rsqrt_seeded, operation by operation: the FP32 seed (numpy's float32 square root) perturbed by seed_err"""
    y = float(np.float32(1.0 / np.sqrt(np.float32(z)))) * (1.0 + seed_err)
    for _ in range(steps):
        y = y * (1.5 - sign * 0.5 * z * y * y)
    return y


def test_newton_reciprocals():
    g = np.random.default_rng(11)
    zs = np.concatenate([10.0 ** g.uniform(-3, 3, 300), [1.0, 0.5, 3.0, 1e-3, 1e3]])
    ulp = 2.0 ** -52
    for e in (2.0 ** -26, -(2.0 ** -26), 0.0):
        worst = max(abs(rcp_newton(float(z), e) * float(z) - 1.0) for z in zs)
        assert worst <= 2 * ulp + e * e, (e, worst)
    # one Newton step squares the seed's error: a 2^-20 seed leaves ~2^-40
    assert abs(rcp_newton(3.0, 2.0 ** -20) * 3.0 - 1.0) < 2.0 ** -39
    zr = np.concatenate([1.0 + 10.0 ** g.uniform(-8, 12, 300), [1.0, 2.0, 1e12]])
    for e in (2.0 ** -22, -(2.0 ** -22), 0.0):
        worst = max(abs(rsqrt_seeded(float(z), e) / float(1 / mpmath.sqrt(mpmath.mpf(float(z)))) - 1.0) for z in zr)
        assert worst <= 4 * ulp, (e, worst)
    # planted: one Newton step short, or a step with the wrong sign
    assert max(abs(rsqrt_seeded(float(z), 2.0 ** -22, steps=1) / float(1 / mpmath.sqrt(mpmath.mpf(float(z)))) - 1.0)
               for z in zr[:20]) > 1e-13
    assert abs(rsqrt_seeded(2.0, 2.0 ** -22, sign=-1.0) * np.sqrt(2.0) - 1.0) > 1e-3


def test_pv2_half_angle():
    g = np.random.default_rng(12)
    ss = np.concatenate([10.0 ** g.uniform(-8, 4, 400), [1e-8, 1.0, 1e4]])
    ulp = 2.0 ** -52
    worst, worst_sqrt = 0.0, 0.0
    for s in ss:
        cb = 1.0 / np.sqrt(1.0 + s * s)
        cbh = np.sqrt((1.0 + cb) / 2.0)
        sbh = abs(s) * cb * (0.5 * (1.0 / cbh))
        exact = mpmath.sin(mpmath.atan(mpmath.mpf(float(s))) / 2)
        worst = max(worst, abs(float(sbh / exact) - 1.0))
        worst_sqrt = max(worst_sqrt, abs(float(np.sqrt((1.0 - cb) / 2.0) / exact) - 1.0))
    assert worst <= 4 * ulp, worst
    assert worst_sqrt > 1e-6  # the cancellation the identity avoids
