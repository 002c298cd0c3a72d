"""brp_hip_search_noalloc against the handle path (brp_hip_create + set_templates + run +
candidates) on the same templates, through the C ABI.

The caller's device memory (the series and the result slots) comes from hipMalloc of the HIP
runtime libbrp_hip.so is linked against, called here through ctypes: the entry point itself
allocates nothing.  Two calls in a row, the second with a smaller series (2 templates at
N = 2^22, then 5 at N = 2^20), each checked slot by slot against the handle path: the same
(bin, level) candidates, powers within 1e-5 relative (bit-equal is expected: the same
kernels run on the same data), no flags; then the size limits' errors.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    """The few HIP runtime calls the test needs as the caller."""

    def __init__(self):
        lib = ctypes.CDLL("libamdhip64.so.7")  # already loaded with libbrp_hip.so: the same runtime
        P = ctypes.POINTER
        lib.hipMalloc.argtypes = [P(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipStreamCreate.argtypes = [P(ctypes.c_void_p)]
        lib.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        lib.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.lib = lib

    def ok(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.lib.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p.value

    def upload(self, dst, arr):
        self.ok(self.lib.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")

    def download(self, arr, src):
        self.ok(self.lib.hipMemcpy(arr.ctypes.data, src, arr.nbytes, D2H), "hipMemcpy D2H")


def _handle_candidates(brp, templates, x):
    count = len(templates)
    with brp.GPU(templates[0], x, 256, 0, count, packed=True, engine=brp.ENGINE_FUSED, group=count) as g:
        g.set_templates(templates)
        g.run()
        return [g.candidates(s) for s in range(count)]


def test_search_noalloc_equals_the_handle_path(brp):
    import workloads

    hip = Hip()
    stream = ctypes.c_void_p()
    hip.ok(hip.lib.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
    total = 0
    try:
        for n, count in ((1 << 22, 2), (1 << 20, 5)):
            w = workloads.bank(count=count, n=n)
            tpl = w.templates
            say(f"[gpu] search_noalloc N={n} templates={count}")
            ref = _handle_candidates(brp, tpl, w.x)
            cap = tpl[0].capacity
            slot_words = 4 + 4 * cap
            series = np.concatenate([w.x, np.zeros(4, np.float32)])
            x_dev = hip.malloc(series.nbytes)
            res_dev = hip.malloc(count * slot_words * 4)
            try:
                hip.upload(x_dev, series)
                brp.search_noalloc(tpl, x_dev, series.size, res_dev, stream.value)
                hip.ok(hip.lib.hipStreamSynchronize(stream), "hipStreamSynchronize")
                words = np.empty(count * slot_words, np.uint32)
                hip.download(words, res_dev)
            finally:
                hip.lib.hipFree(x_dev)
                hip.lib.hipFree(res_dev)
            for s in range(count):
                got, found, flags = brp.slot_candidates(words, cap, s)
                assert flags == 0 and found == len(got), (n, s, flags, found)
                want = ref[s]
                assert np.array_equal(got["bin"], want["bin"]) and np.array_equal(got["level"], want["level"]), (n, s)
                assert np.allclose(got["power"], want["power"], rtol=1e-5, atol=0), (n, s)
                say(f"[gpu]   slot {s}: {found} candidates, bit-equal powers: "
                    f"{bool(np.array_equal(got['power'], want['power']))}")
                total += found
        assert total > 0
        # Limits: more templates than the scratch holds at 2^22, more than 8 per call.
        w = workloads.bank(count=9, n=1 << 22)
        with pytest.raises(RuntimeError, match="bytes of scratch"):
            brp.search_noalloc(w.templates[:3], 1, w.x.size + 4, 1, stream.value)
        with pytest.raises(RuntimeError, match="templates per call"):
            brp.search_noalloc(w.templates[:9], 1, w.x.size + 4, 1, stream.value)
        with pytest.raises(RuntimeError, match="unpadded \\+ 2"):
            brp.search_noalloc(w.templates[:1], 1, w.x.size, 1, stream.value)
    finally:
        hip.lib.hipStreamDestroy(stream)
