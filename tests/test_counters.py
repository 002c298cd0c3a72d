"""In-process hardware counters (profile_counters, src/vulkan.cpp:346-406) over rocprofiler-sdk.

The counters have to be armed before the ROCm runtime starts, so the GPU case runs in a child
process with BRP_HIP_COUNTERS=1; the CPU case checks the unarmed answer."""
import ctypes
import json
import os
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1] + "/boinc-einstein-bpr4-v3d_amd")
import brp_hip as brp, workloads
w = workloads.plain_small()
p = w.templates[0]
out = {}
for name, kw in (("fused", dict(packed=True, engine=brp.ENGINE_FUSED, graph=True)),
                 ("plain", dict(stockham=True, packed=True, radix8=True, fused_power=True))):
    with brp.GPU(p, w.x, 256, 0, 1, **kw) as g:
        g.set_template(p)
        g.run()
        out[name] = json.loads(g.profile_counters())
print("COUNTERS " + json.dumps(out), flush=True)
"""


def test_counters_unarmed_say_why(brp):
    """Without BRP_HIP_COUNTERS=1 at load time the answer names the missing switch (no device call)."""
    if os.environ.get("BRP_HIP_COUNTERS") == "1":
        pytest.skip("counters armed in this process")
    buf = ctypes.create_string_buffer(4096)
    n = brp.load().brp_hip_profile_counters(None, buf, len(buf))
    doc = json.loads(buf.value.decode())
    assert n == len(buf.value)
    assert doc["available"] is False and "BRP_HIP_COUNTERS=1" in doc["reason"]


@pytest.mark.gpu
def test_counters_in_process():
    env = dict(os.environ, BRP_HIP_COUNTERS="1")
    print("[gpu] profile_counters in a child process with BRP_HIP_COUNTERS=1", flush=True)
    proc = subprocess.run([sys.executable, "-c", CHILD, str(ROOT)], env=env, capture_output=True, text=True,
                          timeout=240)
    line = next((l for l in proc.stdout.splitlines() if l.startswith("COUNTERS ")), None)
    assert proc.returncode == 0 and line, proc.stdout[-2000:] + proc.stderr[-2000:]
    out = json.loads(line[len("COUNTERS "):])
    fused, plain = out["fused"], out["plain"]
    assert fused["available"] and plain["available"], out
    by_stage = {s["stage"]: s for s in fused["stages"]}
    for stage in ("pass1_resample_prepare_fft", "pass2_fft_power", "harmonic_candidates", "calibration_copy"):
        assert stage in by_stage, sorted(by_stage)
    p1 = by_stage["pass1_resample_prepare_fft"]
    # N = 65536: C = 256, R = 128, 16 rows per 256-thread block -> 8 blocks of 4 waves
    assert p1["counters"]["SQ_WAVES"] == 32
    assert p1["counters"]["SQ_INSTS_VALU"] > 0 and p1["derived"]["valu_per_wave"] > 100
    # pass 1 writes the exchange: M complex = 4 N bytes
    assert 0.8 * 4 * 65536 <= p1["derived"]["write_bytes"] <= 2.0 * 4 * 65536
    cal = fused["calibration"]
    assert cal["fetch_bytes_per_unit"] > 0 and cal["write_bytes_per_unit"] > 0
    stages = {s["stage"] for s in plain["stages"]}
    assert {"resample_prepare", "fft", "power", "harmonic_candidates"} <= stages, stages
