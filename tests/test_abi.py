"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every
entry point include/gwat_hip.h declares (no compute without a GPU)."""
import ctypes
import os
import re

import gw_analysis_tools_amd as G

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


def test_params_layout_matches_header():
    # 19 doubles then 4 int32, no padding
    assert ctypes.sizeof(G.Params) == 19 * 8 + 4 * 4
    assert G.Params.shift_time.offset == 19 * 8


def test_version_string_without_gpu():
    L = G.lib()
    assert b"gfx950" in L.gwh_version()
