"""ctypes binding of oracle/liboracle.so -- the CPU restatement used as the checker.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The product (libbrp_hip.so and brp_hip.py) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
LIBRARY = ORACLE_DIR / "liboracle.so"


class OParams(ctypes.Structure):
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


class OCandidate(ctypes.Structure):
    _fields_ = [("bin", ctypes.c_uint32), ("level", ctypes.c_uint32), ("power", ctypes.c_float),
                ("reserved", ctypes.c_uint32)]


class ORow(ctypes.Structure):
    _fields_ = [("power", ctypes.c_double), ("period", ctypes.c_double), ("tau", ctypes.c_double),
                ("psi", ctypes.c_double), ("significance", ctypes.c_double), ("harmonics", ctypes.c_uint32),
                ("bin", ctypes.c_uint32)]


CAND_DTYPE = np.dtype([("bin", "<u4"), ("level", "<u4"), ("power", "<f4"), ("reserved", "<u4")])
_lib = None


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-f", str(ORACLE_DIR / "Makefile")], check=True)
    return LIBRARY


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        build()
    lib = ctypes.CDLL(str(LIBRARY))
    P = ctypes.POINTER
    u32, f32, f64, sz, i32 = ctypes.c_uint32, ctypes.c_float, ctypes.c_double, ctypes.c_size_t, ctypes.c_int
    fp, dp = P(ctypes.c_float), P(ctypes.c_double)
    sig = {
        "oracle_last_error": (ctypes.c_char_p, []),
        "oracle_lut_table": (None, [fp]),
        "oracle_lut_sin": (f32, [f32]),
        "oracle_fused_sin": (f32, [f32]),
        "oracle_sample_offset": (f32, [u32, P(OParams)]),
        "oracle_resampling_steps": (i32, [P(OParams)]),
        "oracle_parameters": (i32, [u32, u32, f32, f32, f32, f32, P(OParams)]),
        "oracle_resample": (i32, [fp, sz, P(OParams), fp, P(ctypes.c_int32)]),
        "oracle_preprocess": (i32, [P(ctypes.c_uint8), sz, u32, ctypes.c_double, u32, u32, ctypes.c_float, fp, fp,
                                    P(u32)]),
        "oracle_preprocess_zap": (i32, [P(ctypes.c_uint8), sz, u32, ctypes.c_double, u32, u32, ctypes.c_float,
                                        dp, sz, ctypes.c_double, ctypes.c_double, fp, fp, P(u32)]),
        "oracle_uniform": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64]),
        "oracle_rfft_f64": (i32, [dp, sz, dp]),
        "oracle_rfft_f32": (i32, [fp, sz, fp]),
        "oracle_power": (i32, [fp, sz, fp, dp]),
        "oracle_harmonic": (i32, [fp, sz, P(OParams), fp]),
        "oracle_harmonic_scatter": (i32, [fp, sz, P(OParams), fp]),
        "oracle_select": (sz, [fp, P(OParams), P(OCandidate), sz]),
        "oracle_reference": (i32, [fp, sz, P(OParams), fp, fp, fp]),
        "oracle_chisq_Q": (f64, [f64, f64]),
        "oracle_chisq_Qinv": (f64, [f64, f64]),
        "oracle_search_parameters": (i32, [u32, u32, f64, f64, f64, f64, f64, f64, u32, P(OParams)]),
        "oracle_toplist_thresholds": (None, [P(ORow), u32, P(OParams)]),
        "oracle_toplist_retain": (i32, [P(ORow), u32, P(OCandidate), sz, f64, f64, f64, P(OParams)]),
        "oracle_toplist_final": (sz, [P(ORow), u32, P(ORow), sz]),
        "oracle_benchmark_cpu": (f64, [fp, sz, P(OParams), sz, u32, u32]),
        "oracle_bank_candidates": (i32, [fp, sz, P(OParams), sz, u32, u32, P(OCandidate), sz, P(u32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        raise RuntimeError(load().oracle_last_error().decode())


def as_oparams(p) -> OParams:
    """Byte copy of any 108-byte Params structure."""
    o = OParams()
    ctypes.memmove(ctypes.byref(o), ctypes.byref(p), ctypes.sizeof(OParams))
    return o


def lut_table() -> np.ndarray:
    out = np.empty(130, np.float32)
    load().oracle_lut_table(_fp(out))
    return out


def parameters(n, unpadded, tau, period, psi, dt) -> OParams:
    p = OParams()
    _check(load().oracle_parameters(n, unpadded, tau, period, psi, dt, ctypes.byref(p)))
    return p


def search_parameters(n, samples, sample_us, orbit, f0=400.0, false_alarm=0.08, window=1000) -> OParams:
    p = OParams()
    period, tau, psi = orbit
    _check(load().oracle_search_parameters(n, samples, sample_us, period, tau, psi, f0, false_alarm, window,
                                           ctypes.byref(p)))
    return p


def resample(x: np.ndarray, p) -> tuple[np.ndarray, np.ndarray]:
    p = as_oparams(p)
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(p.n, np.float32)
    idx = np.empty(p.steps, np.int32)
    _check(load().oracle_resample(_fp(x), x.size, ctypes.byref(p), _fp(out),
                                  idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return out, idx


def rfft64(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(2 * (x.size // 2 + 1), np.float64)
    _check(load().oracle_rfft_f64(_dp(x), x.size, _dp(out)))
    return out.view(np.complex128)


def rfft32(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(2 * (x.size // 2 + 1), np.float32)
    _check(load().oracle_rfft_f32(_fp(x), x.size, _fp(out)))
    return out.view(np.complex64)


def power(resampled: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    r = np.ascontiguousarray(resampled, np.float32)
    p32 = np.empty(r.size // 2, np.float32)
    p64 = np.empty(r.size // 2, np.float64)
    _check(load().oracle_power(_fp(r), r.size, _fp(p32), _dp(p64)))
    return p32, p64


def harmonic(power32: np.ndarray, p, scatter: bool = False) -> np.ndarray:
    p = as_oparams(p)
    pw = np.ascontiguousarray(power32, np.float32)
    out = np.empty(5 * p.fundamental, np.float32)
    fn = load().oracle_harmonic_scatter if scatter else load().oracle_harmonic
    _check(fn(_fp(pw), pw.size, ctypes.byref(p), _fp(out)))
    return out


def select(summed: np.ndarray, p) -> np.ndarray:
    p = as_oparams(p)
    s = np.ascontiguousarray(summed, np.float32)
    n = load().oracle_select(_fp(s), ctypes.byref(p), None, 0)
    out = (OCandidate * max(n, 1))()
    load().oracle_select(_fp(s), ctypes.byref(p), out, n)
    return np.frombuffer(bytes(out), dtype=CAND_DTYPE)[:n].copy()


def reference(x: np.ndarray, p):
    """src/science.cpp:113-120: (resampled, index, power32, power64, summed, candidates)."""
    res, idx = resample(x, p)
    p32, p64 = power(res)
    summed = harmonic(p32, p)
    return res, idx, p32, p64, summed, select(summed, p)


def chisq_Q(x, dof):
    return load().oracle_chisq_Q(x, dof)


def chisq_Qinv(q, dof):
    return load().oracle_chisq_Qinv(q, dof)


class OToplist:
    def __init__(self, rows=100):
        self.rows = rows
        self.state = (ORow * (5 * rows))()

    def thresholds(self, p):
        load().oracle_toplist_thresholds(self.state, self.rows, ctypes.byref(p))
        return p

    def retain(self, cands: np.ndarray, orbit, p):
        arr = (OCandidate * max(len(cands), 1))()
        for i, c in enumerate(cands):
            arr[i] = OCandidate(int(c["bin"]), int(c["level"]), float(c["power"]), 0)
        _check(load().oracle_toplist_retain(self.state, self.rows, arr, len(cands), orbit[0], orbit[1], orbit[2],
                                            ctypes.byref(as_oparams(p))))

    def final(self):
        out = (ORow * self.rows)()
        n = load().oracle_toplist_final(self.state, self.rows, out, self.rows)
        return list(out[:n])


def worker_threads() -> int:
    """Host threads for the oracle's thread pool: the affinity set, at most 16 (the host
    share of one GPU)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def bank_candidates(x: np.ndarray, templates, threads: int = 0, cap: int = 8192, fft32: bool = False) -> list:
    """Per-template candidate lists (CAND_DTYPE, (level, bin) order) of the bank loop's
    per-template pipeline with each template's own thresholds, on a thread pool; fft32: the
    CPU baseline's FP32 FFT and power instead of the FP64 ones."""
    arr = (OParams * len(templates))(*[as_oparams(t) for t in templates])
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(len(templates) * cap, dtype=CAND_DTYPE)
    counts = np.zeros(len(templates), dtype=np.uint32)
    _check(load().oracle_bank_candidates(_fp(x), x.size, arr, len(templates), threads or worker_threads(), int(fft32),
                                         out.ctypes.data_as(ctypes.POINTER(OCandidate)), cap,
                                         counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
    return [out[t * cap:t * cap + int(counts[t])].copy() for t in range(len(templates))]


def bank_toplists(templates, orbits, candidates, rows_list=(100,)) -> list:
    """src/workunit_main.cpp:77-136 one template at a time: for each toplist size, raise each
    template's thresholds from the toplist as it stands (src/workload.cpp:126) and retain its
    candidates (src/workload.cpp:129-141), in template order."""
    tops = [OToplist(rows) for rows in rows_list]
    for p, o, c in zip(templates, orbits, candidates):
        for top in tops:
            q = as_oparams(p)
            top.thresholds(q)
            top.retain(c, o, q)
    return tops


def benchmark_cpu(x: np.ndarray, templates, threads: int, reps: int) -> float:
    arr = (OParams * len(templates))(*[as_oparams(t) for t in templates])
    x = np.ascontiguousarray(x, np.float32)
    return load().oracle_benchmark_cpu(_fp(x), x.size, arr, len(templates), threads, reps)


def preprocess(block: np.ndarray, samples: int, bits: int, scale: float, n: int, window: int = 1000,
               rfi_ratio: float = 30.0, zap=None, sample_us: float = 0.0, padding: float = 1.0):
    """src/workload.cpp:29-95 restated in FP64: (preprocessed, decoded, rfi bins replaced);
    zap: optional (lower, upper) Hz ranges."""
    b = np.ascontiguousarray(block, dtype=np.uint8)
    out = np.empty(samples, dtype=np.float32)
    dec = np.empty(samples, dtype=np.float32)
    rfi = ctypes.c_uint32()
    z = np.ascontiguousarray(np.asarray(zap if zap is not None else [], dtype=np.float64).reshape(-1))
    _check(load().oracle_preprocess_zap(b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), samples, bits, scale, n,
                                        window, rfi_ratio, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        z.size // 2, sample_us, padding, _fp(out), _fp(dec), ctypes.byref(rfi)))
    return out, dec, rfi.value


def uniform(seed: int, counter: int) -> float:
    return load().oracle_uniform(seed, counter)

