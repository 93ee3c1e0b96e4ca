"""Host code under the compiler's sanitizers (CPU only):

* the oracle (every model, likelihood, integration rule and Fisher basis) under the address and
  undefined-behaviour sanitizers (`make -C oracle sanitize`);
* the single-call coalescer (gwh_coalesce.h: the lock, queue and lane hand-over the engine's reference-signature
  calls go through) under the thread sanitizer, and under the address / undefined-behaviour sanitizers, with 48
  threads and a host-only stand-in for the launch;
* the Gauss-Legendre rule builder (gwh_quadrature.cpp) under the address / undefined-behaviour sanitizers.

A compiler without the sanitizer runtimes skips; any finding fails."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gw_analysis_tools_amd", "csrc")
HOST = os.path.join(ROOT, "tests", "host")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="no host compiler")


def _build(tmp_path, name, sources, sanitize, includes):
    exe = str(tmp_path / name)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-fno-omit-frame-pointer", "-fsanitize=" + sanitize,
           "-fno-sanitize-recover=all"] + ["-I" + i for i in includes] + sources + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        if "cannot find" in r.stderr and "san" in r.stderr:
            pytest.skip("sanitizer runtime not installed")
        raise AssertionError(r.stderr)
    return exe


def _run(exe, args=(), env=None):
    r = subprocess.run([exe] + list(args), capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, **(env or {})))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    return r.stdout


def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True,
                       text=True, timeout=600)
    if r.returncode != 0 and "san" in r.stderr and "cannot find" in r.stderr:
        pytest.skip("sanitizer runtime not installed")
    assert r.returncode == 0, r.stderr
    out = _run(os.path.join(ROOT, "oracle", "build", "gwo_sanitize"), env={"ASAN_OPTIONS": "detect_leaks=1"})
    assert "run clean" in out


@pytest.mark.parametrize("sanitize", ["thread", "address,undefined"])
def test_coalescer_under_sanitizers(tmp_path, sanitize):
    exe = _build(tmp_path, "coalesce_" + sanitize.replace(",", "_"), [os.path.join(HOST, "coalesce_check.cpp")],
                 sanitize, [CSRC])
    out = _run(exe, ["48", "300"], env={"TSAN_OPTIONS": "halt_on_error=1"})
    print(out.strip())
    assert "0 failures" in out


def test_gauss_legendre_builder_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    exe = _build(tmp_path, "quadrature", [os.path.join(HOST, "quadrature_check.cpp"),
                                          os.path.join(CSRC, "gwh_quadrature.cpp")],
                 "address,undefined", [os.path.join(ROOT, "include")])
    assert "0 failures" in _run(exe)
