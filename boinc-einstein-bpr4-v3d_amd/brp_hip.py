"""Python mirror of the BRP4 per-template search interface over libbrp_hip.so.

The names follow the reference's C++ interface: ``Params`` / ``Candidate``
(include/brp/science.hpp:8-18), ``parameters`` (src/science.cpp:46-65),
``search_parameters`` (src/workload.cpp:108-125), the ``GPU`` class
(include/brp/vulkan.hpp:6-29) and ``Toplist`` (include/brp/workload.hpp:16-23).
Errors raise ``RuntimeError`` with the library's message, as the reference throws
``std::runtime_error``.  The library is required: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import os
import pathlib
from dataclasses import dataclass

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIBRARY = _HERE / "libbrp_hip.so"


class Params(ctypes.Structure):
    """include/brp/science.hpp:8-17 (108 bytes)."""

    _fields_ = [
        ("n", ctypes.c_uint32), ("unpadded", ctypes.c_uint32), ("steps", ctypes.c_uint32), ("logn", ctypes.c_uint32),
        ("tau", ctypes.c_float), ("omega", ctypes.c_float), ("psi", ctypes.c_float), ("dt", ctypes.c_float),
        ("step_inv", ctypes.c_float), ("s0", ctypes.c_float),
        ("fundamental", ctypes.c_uint32), ("harmonic", ctypes.c_uint32), ("window", ctypes.c_uint32),
        ("stage", ctypes.c_uint32), ("threshold", ctypes.c_float * 5), ("capacity", ctypes.c_uint32),
        ("fft_n", ctypes.c_uint32), ("fft_base", ctypes.c_uint32), ("fft_radix", ctypes.c_uint32),
        ("flags", ctypes.c_uint32), ("twiddle3", ctypes.c_uint32), ("twiddle_real", ctypes.c_uint32),
        ("result_slot", ctypes.c_uint32),
    ]

    def copy(self) -> "Params":
        out = Params()
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(Params))
        return out


class Candidate(ctypes.Structure):
    """include/brp/science.hpp:18."""

    _fields_ = [("bin", ctypes.c_uint32), ("level", ctypes.c_uint32), ("power", ctypes.c_float),
                ("reserved", ctypes.c_uint32)]


class Orbit(ctypes.Structure):
    _fields_ = [("period", ctypes.c_double), ("tau", ctypes.c_double), ("psi", ctypes.c_double)]


class BankRow(ctypes.Structure):
    _fields_ = [("power", ctypes.c_double), ("period", ctypes.c_double), ("tau", ctypes.c_double),
                ("psi", ctypes.c_double), ("significance", ctypes.c_double), ("harmonics", ctypes.c_uint32),
                ("bin", ctypes.c_uint32)]


class Options(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint32) for name in (
        "workgroup", "tile", "batch", "software", "stockham", "radix4", "packed", "scatter", "radix8",
        "fused_power", "radix8_store", "engine", "graph", "snapshot", "fp64_phase", "split_c", "precision",
        "group", "split_s1", "split_s2", "lanes", "untimed", "timing_stride", "split_t1", "split_t2")]


class Timings(ctypes.Structure):
    _fields_ = [("wall_ms", ctypes.c_double), ("cpu_ms", ctypes.c_double), ("submit_ms", ctypes.c_double),
                ("stage_ms", ctypes.c_double * 4)]


assert ctypes.sizeof(Params) == 108
assert ctypes.sizeof(Candidate) == 16

ENGINE_PLAIN = 0
ENGINE_FUSED = 1
CANDIDATE_DTYPE = np.dtype([("bin", "<u4"), ("level", "<u4"), ("power", "<f4"), ("reserved", "<u4")])

_lib = None


def load() -> ctypes.CDLL:
    """Load libbrp_hip.so (built in place by build.py); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        raise RuntimeError(f"{LIBRARY} is missing: build it with __graft_entry__.build() first")
    lib = ctypes.CDLL(str(LIBRARY))
    P = ctypes.POINTER
    u32, f32, f64, sz, i32 = ctypes.c_uint32, ctypes.c_float, ctypes.c_double, ctypes.c_size_t, ctypes.c_int
    fp = P(ctypes.c_float)
    sig = {
        "brp_last_error": (ctypes.c_char_p, []),
        "brp_lut_sin": (f32, [f32]),
        "brp_sample_offset": (f32, [u32, P(Params)]),
        "brp_resampling_steps": (i32, [P(Params)]),
        "brp_parameters": (i32, [u32, u32, f32, f32, f32, f32, P(Params)]),
        "brp_search_parameters": (i32, [u32, u32, f64, P(Orbit), f64, f64, u32, P(Params)]),
        "brp_toplist_create": (i32, [u32, P(ctypes.c_void_p)]),
        "brp_toplist_destroy": (None, [ctypes.c_void_p]),
        "brp_toplist_thresholds": (None, [ctypes.c_void_p, P(Params)]),
        "brp_toplist_retain": (i32, [ctypes.c_void_p, P(Candidate), sz, P(Orbit), P(Params)]),
        "brp_toplist_final": (sz, [ctypes.c_void_p, P(BankRow), sz]),
        "brp_toplist_rows": (sz, [ctypes.c_void_p, P(BankRow), sz]),
        "brp_write_candidates": (i32, [ctypes.c_char_p, P(BankRow), sz, f64]),
        "brp_synthetic_series": (i32, [u32, f64, ctypes.c_uint64, f64, f64, f64, f64, f64, P(Orbit), fp]),
        "brp_synthetic_bank": (i32, [ctypes.c_uint64, u32, f64, f64, f64, P(Orbit)]),
        "brp_synthetic_dd": (i32, [u32, f64, ctypes.c_uint64, f64, f64, u32, f64, f64, f64, P(Orbit), u32, u32, f64,
                                   P(ctypes.c_uint8), P(u32)]),
        "brp_hip_preprocess": (i32, [P(ctypes.c_uint8), sz, u32, f64, u32, u32, f32, fp, fp, P(u32)]),
        "brp_hip_bank_search_device": (i32, [ctypes.c_void_p, P(Params), P(Orbit), sz, P(ctypes.c_void_p), sz,
                                             P(u32), P(f64)]),
        "brp_hip_toplist_retain_device": (i32, [P(Params), P(Orbit), sz, P(Candidate), P(u32), ctypes.c_void_p]),
        "brp_hip_preprocess_zap": (i32, [P(ctypes.c_uint8), sz, u32, f64, u32, u32, f32, P(f64), sz, f64, f64, fp,
                                         fp, P(u32)]),
        "brp_uniform": (f64, [ctypes.c_uint64, ctypes.c_uint64]),
        "brp_hip_create": (i32, [P(Params), fp, sz, P(Options), P(ctypes.c_void_p)]),
        "brp_hip_destroy": (i32, [ctypes.c_void_p]),
        "brp_hip_set_templates": (i32, [ctypes.c_void_p, P(Params), sz]),
        "brp_hip_run": (i32, [ctypes.c_void_p, P(Timings), P(ctypes.c_double)]),
        "brp_hip_candidates": (i32, [ctypes.c_void_p, u32, P(Candidate), sz, P(sz)]),
        "brp_hip_snapshot": (i32, [ctypes.c_void_p, u32, fp, P(ctypes.c_int32), fp, fp, fp]),
        "brp_hip_device_info": (sz, [ctypes.c_void_p, ctypes.c_char_p, sz]),
        "brp_hip_allocated_bytes": (ctypes.c_uint64, [ctypes.c_void_p]),
        "brp_hip_profile_counters": (sz, [ctypes.c_void_p, ctypes.c_char_p, sz]),
        "brp_hip_kernel_ms": (i32, [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_uint64), P(ctypes.c_uint64)]),
        "brp_hip_copy_bandwidth": (i32, [sz, i32, P(ctypes.c_double)]),
        "brp_hip_add_series": (i32, [ctypes.c_void_p, P(Params), fp, sz]),
        "brp_hip_bank_search_templates": (i32, [ctypes.c_void_p, P(Params), P(Orbit), sz, P(ctypes.c_void_p), sz,
                                                P(u32), P(f64)]),
        "brp_hip_autotune": (i32, [P(Params), sz, fp, sz, P(Options), u32, P(Options), ctypes.c_char_p, sz]),
        "brp_hip_bank_search": (i32, [ctypes.c_void_p, P(Orbit), sz, u32, f64, f64, f64, u32, ctypes.c_void_p,
                                      P(Params), P(u32), P(ctypes.c_double)]),
        "brp_hip_search_noalloc": (i32, [P(Params), sz, ctypes.c_void_p, sz, ctypes.c_void_p, ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = load().brp_last_error().decode(errors="replace")
        raise RuntimeError(msg or f"brp_hip error {rc}")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


# ------------------------------------------------------------------ science --
def lut_sin(x: float) -> float:
    return load().brp_lut_sin(x)


def sample_offset(i: int, p: Params) -> float:
    return load().brp_sample_offset(i, ctypes.byref(p))


def resampling_steps(p: Params) -> Params:
    _check(load().brp_resampling_steps(ctypes.byref(p)))
    return p


def parameters(n: int, unpadded: int, tau: float, period: float, psi: float, dt: float) -> Params:
    """Bench-semantics template (src/science.cpp:46-65)."""
    p = Params()
    _check(load().brp_parameters(n, unpadded, tau, period, psi, dt, ctypes.byref(p)))
    return p


def search_parameters(n: int, samples: int, sample_us: float, orbit: "Orbit | tuple", f0: float = 400,
                      false_alarm: float = 0.08, window: int = 1000) -> Params:
    """Bank-semantics template (src/workload.cpp:108-125)."""
    o = orbit if isinstance(orbit, Orbit) else Orbit(*orbit)
    p = Params()
    _check(load().brp_search_parameters(n, samples, sample_us, ctypes.byref(o), f0, false_alarm, window, ctypes.byref(p)))
    return p


def synthetic_series(count: int, dt: float, seed: int, mean: float = 0.0, sigma: float = 1.0, spin_hz: float = 0.0,
                     duty: float = 0.05, amplitude: float = 0.0, orbit: "Orbit | tuple | None" = None) -> np.ndarray:
    out = np.empty(count, dtype=np.float32)
    o = None if orbit is None else (orbit if isinstance(orbit, Orbit) else Orbit(*orbit))
    _check(load().brp_synthetic_series(count, dt, seed, mean, sigma, spin_hz, duty, amplitude,
                                       ctypes.byref(o) if o is not None else None, _fptr(out)))
    return out


def synthetic_dd(samples: int, dt: float, seed: int, mean: float = 7.5, sigma: float = 2.0, bits: int = 4,
                 spin_hz: float = 0.0, duty: float = 0.05, amplitude: float = 0.0, orbit=None, n: int = 0,
                 rfi_count: int = 0, rfi_sigma: float = 50.0):
    """DD sample block (4-bit packed, high nibble first, or signed 8-bit); returns (bytes, rfi bins)."""
    n = n or samples
    out = np.zeros(samples // 2 if bits == 4 else samples, dtype=np.uint8)
    bins = np.zeros(max(rfi_count, 1), dtype=np.uint32)
    o = None if orbit is None else (orbit if isinstance(orbit, Orbit) else Orbit(*orbit))
    _check(load().brp_synthetic_dd(samples, dt, seed, mean, sigma, bits, spin_hz, duty, amplitude,
                                   ctypes.byref(o) if o is not None else None, n, rfi_count, rfi_sigma,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                   bins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
    return out, bins[:rfi_count].copy()


def preprocess(block: np.ndarray, samples: int, bits: int, scale: float, n: int, window: int = 1000,
               rfi_ratio: float = 30.0, zap=None, sample_us: float = 0.0, padding: float = 1.0):
    """src/workload.cpp:29-95 on the device: returns (preprocessed, decoded, rfi bins replaced).
    zap (optional): (lower, upper) Hz ranges replaced by noise after whitening, sized by
    sample_us and padding as src/workload.cpp:76-85."""
    b = np.ascontiguousarray(block, dtype=np.uint8)
    out = np.empty(samples, dtype=np.float32)
    dec = np.empty(samples, dtype=np.float32)
    rfi = ctypes.c_uint32()
    z = np.ascontiguousarray(np.asarray(zap if zap is not None else [], dtype=np.float64).reshape(-1))
    _check(load().brp_hip_preprocess_zap(b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), samples, bits, scale, n,
                                         window, rfi_ratio, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         z.size // 2, sample_us, padding, _fptr(out), _fptr(dec),
                                         ctypes.byref(rfi)))
    return out, dec, rfi.value


def synthetic_bank(seed: int, count: int, pmin: float = 660.0, pmax: float = 10000.0, tau_max: float = 0.3):
    arr = (Orbit * count)()
    _check(load().brp_synthetic_bank(seed, count, pmin, pmax, tau_max, arr))
    return arr


def autotune(templates, x: np.ndarray, batch_options: "dict | None" = None, reps: int = 3):
    """Times every compiled four-step split and templates-per-launch value on these
    templates; returns (best option dict, list of candidate reports)."""
    import json
    arr = (Params * len(templates))(*templates)
    xs = np.ascontiguousarray(x, dtype=np.float32)
    kw = dict(workgroup=256, batch=len(templates), packed=1, engine=ENGINE_FUSED, graph=1)
    kw.update(batch_options or {})
    base = Options(**{k: v for k, v in kw.items() if k in dict(Options._fields_)})
    best = Options()
    buf = ctypes.create_string_buffer(1 << 16)
    _check(load().brp_hip_autotune(arr, len(arr), _fptr(xs), xs.size, ctypes.byref(base), reps, ctypes.byref(best),
                                   buf, len(buf)))
    report = [json.loads(line) for line in buf.value.decode().splitlines() if line.strip()]
    return {"split_c": best.split_c, "split_s1": best.split_s1, "split_s2": best.split_s2,
            "split_t1": best.split_t1, "split_t2": best.split_t2, "group": best.group, "lanes": best.lanes}, report


NOALLOC_SCRATCH_BYTES = 64 << 20
NOALLOC_MAX_TEMPLATES = 8


def search_noalloc(templates, input_ptr: int, input_count: int, results_ptr: int, stream: int = 0) -> None:
    """brp_hip_search_noalloc: the fused pipeline for up to 8 templates of one series with
    its buffers in the library's static device scratch; input_ptr / results_ptr are device
    addresses (input_count floats readable, count slots of 16 + 16 * capacity bytes), stream
    a hipStream_t handle (0: default stream).  Only enqueues: synchronize before reading."""
    arr = (Params * len(templates))(*templates)
    _check(load().brp_hip_search_noalloc(arr, len(arr), ctypes.c_void_p(input_ptr), input_count,
                                         ctypes.c_void_p(results_ptr), ctypes.c_void_p(stream)))


def slot_candidates(words: np.ndarray, capacity: int, slot: int) -> tuple[np.ndarray, int, int]:
    """(candidates sorted by (level, bin), count, flags) of result slot `slot` in a uint32
    array of slots of 4 + 4 * capacity words ({count, flags, 0, 0}, then candidates)."""
    rec = words[slot * (4 + 4 * capacity):(slot + 1) * (4 + 4 * capacity)]
    count, flags = int(rec[0]), int(rec[1])
    c = rec[4:4 + 4 * min(count, capacity)].view(CANDIDATE_DTYPE).copy()
    c.sort(order=["level", "bin"])
    return c, count, flags


def copy_bandwidth(nbytes: int, reps: int = 10) -> float:
    """GB/s (read + write) of a plain device-to-device streaming copy of nbytes."""
    out = ctypes.c_double()
    _check(load().brp_hip_copy_bandwidth(nbytes, reps, ctypes.byref(out)))
    return out.value


def uniform(seed: int, counter: int) -> float:
    return load().brp_uniform(seed, counter)


def candidates_array(c) -> np.ndarray:
    """Structured array (bin, level, power, reserved) from a sequence of Candidate."""
    out = np.zeros(len(c), CANDIDATE_DTYPE)
    for i, x in enumerate(c):
        out[i] = (x.bin, x.level, x.power, x.reserved)
    return out


# ----------------------------------------------------------------- toplist --
class Toplist:
    """src/workload.cpp:126-159 with `rows` rows per level (reference: 100)."""

    def __init__(self, rows: int = 100):
        self._h = ctypes.c_void_p()
        _check(load().brp_toplist_create(rows, ctypes.byref(self._h)))
        self.rows = rows

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.brp_toplist_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    def thresholds(self, p: Params) -> Params:
        load().brp_toplist_thresholds(self._h, ctypes.byref(p))
        return p

    def retain(self, candidates, orbit, p: Params) -> None:
        o = orbit if isinstance(orbit, Orbit) else Orbit(*orbit)
        arr = (Candidate * len(candidates))(*candidates) if not isinstance(candidates, ctypes.Array) else candidates
        _check(load().brp_toplist_retain(self._h, arr, len(arr), ctypes.byref(o), ctypes.byref(p)))

    def final(self):
        out = (BankRow * self.rows)()
        n = load().brp_toplist_final(self._h, out, self.rows)
        return list(out[:n])

    def table(self):
        out = (BankRow * (5 * self.rows))()
        load().brp_toplist_rows(self._h, out, 5 * self.rows)
        return list(out)


def toplist_retain_device(top: "Toplist", templates, orbits, candidate_lists) -> None:
    """Toplist::retain of each template's candidate list on the device, in order."""
    tarr = (Params * len(templates))(*templates)
    oarr = (Orbit * len(orbits))(*[o if isinstance(o, Orbit) else Orbit(*o) for o in orbits])
    flat = [c for lst in candidate_lists for c in lst]
    carr = (Candidate * max(len(flat), 1))(*flat)
    counts = (ctypes.c_uint32 * max(len(candidate_lists), 1))(*[len(lst) for lst in candidate_lists])
    _check(load().brp_hip_toplist_retain_device(tarr, oarr, len(tarr), carr, counts, top.handle))


def write_candidates(path: str, rows, seconds: float) -> None:
    arr = (BankRow * len(rows))(*rows)
    _check(load().brp_write_candidates(os.fsencode(path), arr, len(rows), seconds))


# --------------------------------------------------------------------- GPU --
@dataclass
class Snapshot:
    resampled: np.ndarray
    index: np.ndarray | None
    fft: np.ndarray
    power: np.ndarray
    summed: np.ndarray
    candidates: np.ndarray


class GPU:
    """include/brp/vulkan.hpp:12-29 over the HIP backend.

    The positional arguments after ``input`` are the reference constructor's; the
    keyword-only ones select the MI355X engine and its options.
    """

    def __init__(self, p: Params, input: np.ndarray, workgroup: int = 256, tile: int = 0, batch: int = 1,
                 software: bool = False, stockham: bool = False, radix4: bool = False, packed: bool = False,
                 scatter: bool = False, radix8: bool = False, fused_power: bool = False, radix8_store: bool = False,
                 *, engine: int = ENGINE_PLAIN, graph: bool = False, snapshot: bool = False, fp64_phase: bool = False,
                 split_c: int = 0, precision: int = 0, group: int = 0, split_s1: int = 0, split_s2: int = 0,
                 lanes: int = 0, untimed: bool = False, timing_stride: int = 0, split_t1: int = 0,
                 split_t2: int = 0):
        lib = load()
        x = np.ascontiguousarray(input, dtype=np.float32)
        o = Options(workgroup, tile, batch, software, stockham, radix4, packed, scatter, radix8, fused_power,
                    radix8_store, engine, graph, snapshot, fp64_phase, split_c, precision, group, split_s1,
                    split_s2, lanes, untimed, timing_stride, split_t1, split_t2)
        self._h = ctypes.c_void_p()
        self.p = p.copy()
        self.batch = batch
        self.engine = engine
        self.count = 1
        _check(lib.brp_hip_create(ctypes.byref(self.p), _fptr(x), x.size, ctypes.byref(o), ctypes.byref(self._h)))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.brp_hip_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_template(self, p: Params) -> None:
        self.set_templates([p])

    def set_templates(self, ps) -> None:
        arr = (Params * len(ps))(*ps)
        _check(load().brp_hip_set_templates(self._h, arr, len(ps)))
        self.count = len(ps)
        self._templates = [q.copy() for q in ps]

    def run(self) -> dict:
        t = Timings()
        per = (ctypes.c_double * (4 * self.count))()
        _check(load().brp_hip_run(self._h, ctypes.byref(t), per))
        return {"wall_ms": t.wall_ms, "cpu_ms": t.cpu_ms, "submit_ms": t.submit_ms, "stage_ms": list(t.stage_ms),
                "template_stage_ms": [list(per[4 * i:4 * i + 4]) for i in range(self.count)]}

    def candidates(self, slot: int | None = None) -> np.ndarray:
        if slot is None:
            slot = self.count - 1
        lib = load()
        n = ctypes.c_size_t()
        _check(lib.brp_hip_candidates(self._h, slot, None, 0, ctypes.byref(n)))
        out = (Candidate * max(n.value, 1))()
        _check(lib.brp_hip_candidates(self._h, slot, out, n.value, ctypes.byref(n)))
        return np.frombuffer(out, dtype=CANDIDATE_DTYPE)[:n.value].copy()

    def snapshot(self, slot: int | None = None, p: Params | None = None) -> Snapshot:
        if slot is None:
            slot = self.count - 1
        q = p or (self._templates[slot] if hasattr(self, "_templates") else self.p)
        n = q.n
        res = np.empty(n, np.float32)
        idx = np.empty(q.steps, np.int32)
        fft = np.empty(2 * (n // 2 + 1), np.float32)
        power = np.empty(n // 2, np.float32)
        summed = np.empty(5 * q.fundamental, np.float32)
        rc = load().brp_hip_snapshot(self._h, slot, _fptr(res), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     _fptr(fft), _fptr(power), _fptr(summed))
        if rc != 0 and b"index snapshot" in load().brp_last_error():
            idx = None
            rc = load().brp_hip_snapshot(self._h, slot, _fptr(res), None, _fptr(fft), _fptr(power), _fptr(summed))
        _check(rc)
        return Snapshot(res, idx, fft.view(np.complex64), power, summed, self.candidates(slot))

    def device_info(self) -> str:
        buf = ctypes.create_string_buffer(1024)
        load().brp_hip_device_info(self._h, buf, 1024)
        return buf.value.decode()

    def allocated_bytes(self) -> int:
        return int(load().brp_hip_allocated_bytes(self._h))

    def profile_counters(self) -> str:
        """Per-kernel hardware counters of the last batch as JSON (collected on each call; in
        process only when BRP_HIP_COUNTERS=1 was set before the library loaded)."""
        cap = 1 << 20
        buf = ctypes.create_string_buffer(cap)
        n = load().brp_hip_profile_counters(self._h, buf, cap)
        if n >= cap:
            raise RuntimeError(f"profile_counters: {n} bytes of JSON exceed the {cap}-byte buffer")
        return buf.value.decode()

    def kernel_ms(self) -> tuple[list[float], int, int]:
        """Per-stage device ms since the previous call (which resets them), the number of
        launch sequences (groups or templates) they cover and the templates in those."""
        out = (ctypes.c_double * 4)()
        n = ctypes.c_uint64()
        m = ctypes.c_uint64()
        _check(load().brp_hip_kernel_ms(self._h, out, ctypes.byref(n), ctypes.byref(m)))
        return list(out), n.value, m.value

    def add_series(self, p: Params, input: np.ndarray) -> None:
        """Registers another observation (fused engine): templates name it by (n, unpadded)."""
        x = np.ascontiguousarray(input, dtype=np.float32)
        _check(load().brp_hip_add_series(self._h, ctypes.byref(p), _fptr(x), x.size))

    def bank_search_templates(self, templates, orbits, tops):
        """Bank loop over prepared templates of any registered series, tops[i] the toplist of
        series i; returns (counts, gpu_ms)."""
        tarr = (Params * len(templates))(*templates)
        oarr = (Orbit * len(orbits))(*[o if isinstance(o, Orbit) else Orbit(*o) for o in orbits])
        harr = (ctypes.c_void_p * len(tops))(*[tp.handle for tp in tops])
        counts = (ctypes.c_uint32 * len(tarr))()
        gpu_ms = ctypes.c_double()
        _check(load().brp_hip_bank_search_templates(self._h, tarr, oarr, len(tarr), harr, len(tops), counts,
                                                    ctypes.byref(gpu_ms)))
        return np.frombuffer(counts, dtype=np.uint32).copy(), gpu_ms.value

    def bank_search_device(self, templates, orbits, tops):
        """bank_search_templates with the toplists retained on the device; returns
        (counts, gpu_ms) and leaves each toplist as the host loop would."""
        tarr = (Params * len(templates))(*templates)
        oarr = (Orbit * len(orbits))(*[o if isinstance(o, Orbit) else Orbit(*o) for o in orbits])
        harr = (ctypes.c_void_p * len(tops))(*[tp.handle for tp in tops])
        counts = (ctypes.c_uint32 * len(tarr))()
        gpu_ms = ctypes.c_double()
        _check(load().brp_hip_bank_search_device(self._h, tarr, oarr, len(tarr), harr, len(tops), counts,
                                                 ctypes.byref(gpu_ms)))
        return np.frombuffer(counts, dtype=np.uint32).copy(), gpu_ms.value

    def bank_search(self, orbits, samples: int, sample_us: float, top: Toplist, f0: float = 400,
                    false_alarm: float = 0.08, window: int = 1000):
        """src/workunit_main.cpp:77-136 in batches of `batch`; returns (templates, counts, gpu_ms)."""
        arr = orbits if isinstance(orbits, ctypes.Array) else (Orbit * len(orbits))(*orbits)
        tpl = (Params * len(arr))()
        counts = (ctypes.c_uint32 * len(arr))()
        gpu_ms = ctypes.c_double()
        _check(load().brp_hip_bank_search(self._h, arr, len(arr), samples, sample_us, f0, false_alarm, window,
                                          top.handle, tpl, counts, ctypes.byref(gpu_ms)))
        return list(tpl), np.frombuffer(counts, dtype=np.uint32).copy(), gpu_ms.value


TWO_PI = 2 * math.pi
