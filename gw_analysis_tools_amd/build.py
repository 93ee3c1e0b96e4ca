"""Build the engine (HIP for gfx950) into gw_analysis_tools_amd/libgwat_hip.so.

    python -m gw_analysis_tools_amd.build          # engine
    python -m gw_analysis_tools_amd.build --all    # engine + CPU oracle (test infrastructure)

hipcc is driven directly (no JIT cache): the built shared object lives in
the package directory so it travels with the source tree.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgwat_hip.so")
ARCH = os.environ.get("GWH_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result"]


def sources():
    csrc = os.path.join(HERE, "csrc")
    return sorted(glob.glob(os.path.join(csrc, "*.hip")) + glob.glob(os.path.join(csrc, "*.cpp")))


def build_engine(verbose=False, jobs=8):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for s in sources():
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [hipcc] + HIPFLAGS + ["-I", os.path.join(ROOT, "include"), "-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd))
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    cmd = [hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + \
        ["-L" + rocm_lib, "-lrocfft", "-Wl,-rpath," + rocm_lib, "-lpthread"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_host_programs(verbose)
    return LIB


def build_host_programs(verbose=False):
    """host programs over the library (host/*.cpp -> package directory), e.g. the threaded single-call driver"""
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    for src in sorted(glob.glob(os.path.join(HERE, "host", "*.cpp"))):
        exe = os.path.join(HERE, os.path.splitext(os.path.basename(src))[0])
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-pthread", "-Wall", "-I", os.path.join(ROOT, "include"),
               src, "-o", exe + ".tmp", "-L" + HERE, "-lgwat_hip", "-Wl,-rpath,$ORIGIN", "-L" + rocm_lib,
               "-Wl,-rpath," + rocm_lib]
        if verbose:
            print(" ".join(cmd))
        subprocess.check_call(cmd)
        os.replace(exe + ".tmp", exe)


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()
                     and "warning: feature flag" not in l)
    if p.returncode != 0:
        sys.stderr.write(text + "\n")
        raise RuntimeError("hipcc failed on " + src)
    if text.strip():
        sys.stderr.write(text + "\n")


def build_oracle():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


if __name__ == "__main__":
    build_engine(verbose="-v" in sys.argv)
    if "--all" in sys.argv:
        build_oracle()
    print(LIB)
