"""The FLOP-equivalent table bench.py prices the per-bin kernels with is the operation count of the reference
algorithm as the oracle restates it (oracle/build/opcount runs the oracle's sources on a counting scalar):
re-derived here and compared with the frozen constants."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_frozen_fe_table_equals_the_operation_count():
    exe = os.path.join(ROOT, "oracle", "build", "opcount")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    t = json.loads(out)
    import bench
    assert t["additive"] is True
    assert tuple(t["amp_region_fe"]) == bench.FE_AMP
    assert tuple(t["phase_region_fe"]) == bench.FE_PHASE
    assert t["common_per_bin_fe"] == bench.FE_COMMON
    assert t["per_detector_per_bin_fe"] == bench.FE_DET
    assert t["imrphenompv2_extra_per_bin_fe"] == bench.FE_PV2_BIN
    assert t["imrphenompv2_per_detector_per_bin_fe"] - t["per_detector_per_bin_fe"] == bench.FE_PV2_DET
    # the region totals the table reproduces
    totals = [bench.FE_COMMON + bench.FE_AMP[a] + bench.FE_PHASE[p] for a, p in ((0, 0), (1, 0), (1, 1), (1, 2),
                                                                                  (2, 2))]
    assert totals == t["measured_region_totals"]
