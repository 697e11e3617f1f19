"""The benchmark's host-side contract (no GPU): one GPU only, and the counter
figures it reports come from a well-formed profile block."""
import json
import os
import re
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no bench line for a refused run


def test_profile_block_is_well_formed():
    """The JSON block bench.py reads its HBM traffic and lane figures from:
    keyed by the hash of the sources it was measured on, with the copy
    calibration and the two kernels the bench line reports."""
    text = open(bench.PMC_PROFILE).read()
    m = re.search(r"```json\n(\{.*?\})\n```", text, re.S)
    assert m, "no JSON block"
    d = json.loads(m.group(1))
    assert re.fullmatch(r"[0-9a-f]{64}", d["source_sha256"])
    assert 1.0 <= d["calibration"]["fetch"] <= 4.0 and 0.5 <= d["calibration"]["write"] <= 2.0
    for k in ("wave_extend", "wave_shade"):
        kp = d["kernels"][k]
        assert kp["hbm_bytes_per_dispatch"] > 0
        assert 0.0 < kp["lanes_active"] <= 1.0 and 0.0 < kp["wave_wait_fraction"] < 1.0
    # the hash is of the sources as they are: bench.py reports the block only then
    assert (bench.measured_profile() is not None) == (d["source_sha256"] == bench.source_hash())
