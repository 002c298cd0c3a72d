"""The C ABI library loads and exports every function include/brp_hip.h declares.

CPU only: no compute call is made; without a GPU the device entry points must fail
with an error code and a message, not crash.
"""
import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "brp_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(brp_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("brp_hip_create", "brp_hip_destroy", "brp_hip_set_templates", "brp_hip_run",
                     "brp_hip_candidates", "brp_hip_snapshot", "brp_hip_device_info", "brp_hip_allocated_bytes",
                     "brp_hip_profile_counters", "brp_parameters", "brp_search_parameters", "brp_toplist_retain"):
        assert required in names


def test_library_exports_every_declared_symbol(brp):
    lib = ctypes.CDLL(str(brp.LIBRARY))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts(brp):
    assert ctypes.sizeof(brp.Params) == 108
    assert ctypes.sizeof(brp.Candidate) == 16
    assert ctypes.sizeof(brp.Options) == 25 * 4
    assert ctypes.sizeof(brp.BankRow) == 48


def test_create_validates_options_before_touching_the_device(brp):
    import numpy as np

    p = brp.parameters(1 << 12, 1 << 12, 0.01, 600.0, 0.5, 64e-6)
    x = np.zeros(p.unpadded, np.float32)
    # src/vulkan.cpp:279-283: option combinations are rejected with the reference's messages.
    with pytest.raises(RuntimeError, match="fused power requires packed"):
        brp.GPU(p, x, 256, 0, 1, fused_power=True)
    with pytest.raises(RuntimeError, match="radix8 requires Stockham"):
        brp.GPU(p, x, 256, 0, 1, radix8=True)
    with pytest.raises(RuntimeError, match="input length mismatch"):
        brp.GPU(p, x[:-1], 256, 0, 1)
    with pytest.raises(RuntimeError, match="invalid workgroup"):
        brp.GPU(p, x, 512, 0, 1)
    # The precision variant: a bit mask of the passes whose radix-16 stages run on MFMA.
    with pytest.raises(RuntimeError, match="unknown precision variant"):
        brp.GPU(p, x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, precision=4)
    with pytest.raises(RuntimeError, match="needs the fused engine"):
        brp.GPU(p, x, 256, 0, 1, packed=True, precision=2)
    with pytest.raises(RuntimeError, match="FP32 index rule only"):
        brp.GPU(p, x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, precision=2, fp64_phase=True)


def test_create_without_device_fails_cleanly(brp):
    import numpy as np

    try:
        import torch  # noqa: F401  (only to know whether this host has a GPU)
        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    p = brp.parameters(1 << 12, 1 << 12, 0.01, 600.0, 0.5, 64e-6)
    x = np.zeros(p.unpadded, np.float32)
    with pytest.raises(RuntimeError):
        brp.GPU(p, x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED)


def test_cpp_mirror_compiles_links_and_runs_host_science(brp, tmp_path):
    """include/brp/hip.hpp (the reference-shaped C++ class and science functions) compiles
    with g++ -std=c++20, links against the library, and its host-only calls agree with the
    Python binding (no GPU call is made)."""
    import shutil
    import subprocess

    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src = tmp_path / "mirror.cpp"
    src.write_text(r'''
#include "brp/hip.hpp"
#include <cstdio>
int main() {
    brp::Params p = brp::parameters(65536, 65536, 0.05f, 3600.0f, 1.0f, 64e-6f);
    brp::Toplist top(100);
    std::printf("%u %u %.9g %.9g\n", p.steps, p.fundamental, double(p.s0), double(brp::lut_sin(1.25f)));
    try {
        brp::parameters(100, 100, 0.01f, 100.0f, 0.1f, 64e-6f);
    } catch (const std::runtime_error& e) {
        std::printf("error: %s\n", e.what());
    }
    return 0;
}
''')
    exe = tmp_path / "mirror"
    lib_dir = brp.LIBRARY.parent
    subprocess.run(["g++", "-std=c++20", f"-I{ROOT / 'include'}", str(src), "-o", str(exe), f"-L{lib_dir}",
                    "-lbrp_hip", f"-Wl,-rpath,{lib_dir}"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    p = brp.parameters(65536, 65536, 0.05, 3600.0, 1.0, 64e-6)
    steps, fundamental, s0, lut = out[0].split()
    assert int(steps) == p.steps and int(fundamental) == p.fundamental
    assert float(s0) == pytest.approx(p.s0, rel=1e-7) and float(lut) == pytest.approx(brp.lut_sin(1.25), rel=1e-7)
    assert out[1].startswith("error: invalid observation/template")


def test_entry_points_reject_null_arguments(brp):
    """Every handle entry point checks its pointers before any device call: a null handle or
    output pointer is an error status with a message, never a crash."""
    import ctypes

    lib = brp.load()
    p = brp.parameters(1 << 12, 1 << 12, 0.01, 600.0, 0.5, 64e-6)
    n = ctypes.c_size_t()
    calls = [
        ("brp_hip_create", lambda: lib.brp_hip_create(None, None, 0, None, None)),
        ("brp_hip_set_templates", lambda: lib.brp_hip_set_templates(None, ctypes.byref(p), 1)),
        ("brp_hip_run", lambda: lib.brp_hip_run(None, None, None)),
        ("brp_hip_candidates", lambda: lib.brp_hip_candidates(None, 0, None, 0, ctypes.byref(n))),
        ("brp_hip_snapshot", lambda: lib.brp_hip_snapshot(None, 0, None, None, None, None, None)),
        ("brp_hip_add_series", lambda: lib.brp_hip_add_series(None, ctypes.byref(p), None, 0)),
        ("brp_hip_bank_search",
         lambda: lib.brp_hip_bank_search(None, None, 0, 0, 0.0, 0.0, 0.0, 0, None, None, None, None)),
        ("brp_hip_bank_search_templates",
         lambda: lib.brp_hip_bank_search_templates(None, None, None, 0, None, 0, None, None)),
        ("brp_hip_bank_search_device",
         lambda: lib.brp_hip_bank_search_device(None, None, None, 0, None, 0, None, None)),
    ]
    for name, call in calls:
        rc = call()
        assert rc != 0, name
        assert lib.brp_last_error(), name
    assert lib.brp_hip_destroy(None) == 0  # like delete of a null pointer


def test_preprocess_validates_before_touching_the_device(brp):
    """brp_hip_preprocess checks the DD metadata, the lengths and the window before any
    device call, with the reference's messages (src/workload.cpp:37-45, 57)."""
    import numpy as np

    block = np.zeros(512, np.uint8)
    cases = [
        (dict(samples=1024, bits=3, n=1024), "unsupported DD sample format"),
        (dict(samples=1023, bits=4, n=1024), "odd packed DD length"),
        (dict(samples=1000, bits=4, n=1000), "observation exceeds supported length"),
        (dict(samples=2048, bits=4, n=1024), "invalid DD metadata"),
        (dict(samples=1024, bits=4, n=1024, window=63), "unsupported whitening window"),
        (dict(samples=1024, bits=4, n=1024, window=1024), "unsupported whitening window"),
        (dict(samples=1024, bits=4, n=1024, rfi_ratio=-1.0), "invalid RFI ratio"),
        (dict(samples=1024, bits=4, n=1024, zap=[(1.0, 2.0)], sample_us=64.0, padding=0.5), "invalid zap arguments"),
    ]
    for kw, message in cases:
        args = dict(window=64, rfi_ratio=0.0)
        args.update(kw)
        with pytest.raises(RuntimeError, match=message):
            brp.preprocess(block, args.pop("samples"), args.pop("bits"), 1.0, args.pop("n"), **args)
