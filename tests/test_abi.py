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
    assert ctypes.sizeof(brp.Options) == 23 * 4
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
