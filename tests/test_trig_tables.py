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
  the wrong sign is caught."""
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
