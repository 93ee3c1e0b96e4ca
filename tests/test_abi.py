"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every
entry point include/gwat_hip.h declares (no compute without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from gw_analysis_tools_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gwh_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(G.LIB_PATH)
    names = declared("gwat_hip.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(G.EXPORTS) == names


def test_params_layout_matches_header(tmp_path):
    """every field of the ctypes / numpy view at the offset the C compiler gives it in include/gwat_hip.h"""
    import shutil
    import subprocess
    assert ctypes.sizeof(G.Params) == G.PARAMS_DTYPE.itemsize
    for name, _ in G.Params._fields_:
        assert getattr(G.Params, name).offset == G.PARAMS_DTYPE.fields[name][1], name
    if shutil.which("gcc") is None:
        return
    names = [n for n, _ in G.Params._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "gwat_hip.h"\nint main(void){\n'
                   'printf("%zu\\n", sizeof(gwh_params));\n' +
                   "".join('printf("%%zu\\n", offsetof(gwh_params, %s));\n' % n for n in names) + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(G.Params)
    assert out[1:] == [getattr(G.Params, n).offset for n in names]


def test_version_string_without_gpu():
    L = G.lib()
    assert b"gfx950" in L.gwh_version()


# ---- host-side helpers of the boundary (pure host code in the library; no GPU needed) ----

def test_detector_names_follow_the_reference_table():
    for name, i in [("Hanford", 0), ("hanford", 0), ("Livingston", 1), ("Virgo", 2), ("Kagra", 3),
                    ("Indigo", 4), ("Cosmic Explorer", 5), ("CE", 5), ("ET1", 6), ("einstein telescope 2", 7),
                    ("ET3", 8)]:
        assert G.detector_id(name) == i, name
    # names the reference's detector_util does not know (it exits there; the engine returns -1)
    for bad in ("KAGRA", "H1", "", "Hanford "):
        assert G.detector_id(bad) == -1


def test_dtoa_and_gmst_match_oracle():
    assert G.gps_to_gmst_radian(1126259462.4) == O.gps_to_gmst_radian(1126259462.4)
    g = np.random.default_rng(5)
    for _ in range(20):
        ra, dec, gm = g.uniform(0, 6.28), g.uniform(-1.5, 1.5), g.uniform(0, 6.28)
        for d1 in range(9):
            for d2 in range(9):
                assert abs(G.dtoa(ra, dec, gm, d1, d2) - O.dtoa(ra, dec, gm, d1, d2)) <= 1e-15
    # antisymmetry and light-travel bound
    assert G.dtoa(1., .3, 2., 0, 1) == -G.dtoa(1., .3, 2., 1, 0)
    assert abs(G.dtoa(1., .3, 2., 0, 1)) < 0.011


@pytest.mark.parametrize("basis,method,dim,sky", [(G.BASIS_PHENOMD_SKYAVG, "IMRPhenomD", 7, 1),
                                                  (G.BASIS_PHENOMD_FULL, "IMRPhenomD", 11, 0),
                                                  (G.BASIS_PHENOMD_MCMC, "MCMC_IMRPhenomD", 11, 0),
                                                  (G.BASIS_PV2_FULL, "IMRPhenomPv2", 13, 0),
                                                  (G.BASIS_PV2_MCMC, "MCMC_IMRPhenomPv2", 15, 0)])
def test_unpack_repack_match_oracle(basis, method, dim, sky):
    kw = dict(W.INJECTION, spin1=[0.2, -0.1, 0.3], spin2=[0.05, 0.3, -0.1], sky_average=sky)
    if basis == G.BASIS_PV2_FULL:
        kw.update(chip=0.4, phip=1.2)
    g, o = G.make_params(**kw), O.make_params(**kw)
    v, lf = G.unpack_parameters(g, basis, dim)
    vo, lfo = O.unpack_parameters(o, method, dim)
    assert np.allclose(v, vo, rtol=1e-15, atol=0) and np.array_equal(lf, lfo)
    v2 = v * (1 + 1e-3 * np.arange(dim) / dim)
    a = G.repack_parameters(v2, basis, g)
    b = O.repack_parameters(v2, method, dim, o)
    for name in ("mass1", "mass2", "DL", "tc", "phiRef", "RA", "DEC", "psi", "incl", "chip", "phip"):
        assert abs(getattr(a, name) - getattr(b, name)) <= 1e-13 * max(1.0, abs(getattr(b, name))), name
    for i in range(3):
        assert abs(a.spin1[i] - b.spin1[i]) <= 1e-13 and abs(a.spin2[i] - b.spin2[i]) <= 1e-13


# ---- the gwatpy C entry points (include/gwatpy_abi.h) ----

def _declared_plain(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*_py(?:_destructor)?)\s*\(", text)))


def test_gwatpy_entry_points_exported():
    L = ctypes.CDLL(G.LIB_PATH)
    names = _declared_plain("gwatpy_abi.h")
    assert "MCMC_likelihood_extrinsic_py" in names and "gen_params_base_py_destructor" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_gwatpy_host_helpers_and_errors():
    from gw_analysis_tools_amd import gwatpy_compat as P
    L = P.lib()
    assert L.gps_to_GMST_radian_py(1126259462.4) == O.gps_to_gmst_radian(1126259462.4)
    assert L.DTOA_DETECTOR_py(1.2, -0.3, 0.7, b"Hanford", b"Virgo") == O.dtoa(1.2, -0.3, 0.7, 0, 2)
    assert np.isnan(L.DTOA_DETECTOR_py(1.2, -0.3, 0.7, b"Hanford", b"H1"))
    f = np.linspace(20, 30, 5)
    # fields the engine does not model are refused at use, before any device work
    for kw in (dict(equatorial_orientation=True), dict(horizon_coord=True), dict(NSflag1=True), dict(Nmod=1),
               dict(Nmod=5, bppe=[-7.] * 5, betappe=[1e-5] * 5)):
        p = P.gen_params(mass1=30, mass2=20, **kw)
        st, _, _ = P.fourier_waveform(f, "IMRPhenomD", p)
        assert st == G.GWH_ERR_UNSUPPORTED, kw
        assert b"does not model" in G.lib().gwh_last_error()
        p.close()
    p = P.gen_params(mass1=30, mass2=20)
    st, _, _ = P.fourier_waveform(f, "IMRPhenomPv3", p)
    assert st == G.GWH_ERR_UNSUPPORTED
    st, _ = P.fourier_detector_response(f, "LIGO", "IMRPhenomD", p)
    assert st == G.GWH_ERR_UNSUPPORTED
    # a foreign pointer is not a handle
    bad = type("H", (), {"handle": ctypes.c_void_p(12345).value})()
    assert L.fourier_waveform_py(f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 5, None, None, None, None,
                                 b"IMRPhenomD", bad.handle) == G.GWH_ERR_INVALID
    # the 'K' letter maps to a name detector_util does not know
    ll = P.mcmc_likelihood_extrinsic(p, "IMRPhenomD", f, [np.zeros(5, complex)], [np.ones(5)], "K")
    assert np.isnan(ll)
    p.close()


def test_gauleg_nodes_and_weights():
    """gwh_gauleg (gauleg, ortho_basis.cpp:14) against numpy's Gauss-Legendre rule mapped to [x1, x2], ascending,
    and exact for polynomials of degree 2n - 1.  The reference stops Newton's method once a step moves a root by
    at most 1e-10 and takes the weight from that step's derivative, so weights next to the ends of a long rule
    (small 1 - z^2) carry up to ~1e-6 relative of that last step; short rules agree to 1e-8."""
    for n, (x1, x2) in ((1, (0., 1.)), (7, (20., 1024.)), (64, (np.log10(20.), np.log10(2048.))), (801, (-1., 1.))):
        x, w = G.gauleg(x1, x2, n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(x, 0.5 * (x2 - x1) * xr + 0.5 * (x2 + x1), rtol=0, atol=1e-12 * max(1, abs(x2)))
        np.testing.assert_allclose(w, 0.5 * (x2 - x1) * wr, rtol=1e-8 if n <= 64 else 2e-6, atol=0)
        assert np.all(np.diff(x) > 0)
        k = 2 * n - 1
        assert abs(np.sum(w * x ** k) - (x2 ** (k + 1) - x1 ** (k + 1)) / (k + 1)) < 1e-9 * max(1, abs(x2) ** (k + 1))
