"""The command-line drivers bin/brp_workunit and bin/brp_bench (src/workunit_main.cpp,
src/main.cpp): argument and file-format errors with the reference's messages and exit
codes (no device needed, they fail before any device call), and on the GPU the end-to-end
workunit against the library's own bank pipeline and the benchmark's checks."""
import gzip
import json
import math
import pathlib
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
BIN = ROOT / "boinc-einstein-bpr4-v3d_amd" / "bin"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def exe(name):
    path = BIN / name
    if not path.exists():
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")
    return str(path)


def run(cmd, timeout=120):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


def write_dd(path, block, samples, sample_us=64.0, scale=1.0, gz=True, extra=b""):
    """A DD file (src/workload.cpp:29-56): 1168-byte header (sample_us double 0,
    observation_s double 1, dm double 13, scale double 14, samples uint32 at byte 128),
    then the sample block."""
    header = bytearray(1168)
    struct.pack_into("<d", header, 0, sample_us)
    struct.pack_into("<d", header, 8, samples * sample_us * 1e-6)
    struct.pack_into("<d", header, 13 * 8, 0.0)
    struct.pack_into("<d", header, 14 * 8, scale)
    struct.pack_into("<I", header, 128, samples)
    data = bytes(header) + bytes(np.asarray(block, dtype=np.uint8)) + extra
    path.write_bytes(gzip.compress(data) if gz else data)
    return path


def workunit_args(tmp, dd, bank, zap, extra=(), preprocessed=None):
    args = [exe("brp_workunit"), "--input", str(dd), "--bank", str(bank), "--zap", str(zap), "--padding", "1",
            "--f0", "400", "--false-alarm", "0.08", "--window", "1000", "--output", str(tmp / "out.txt"),
            "--metrics", str(tmp / "metrics.jsonl"), *extra]
    if preprocessed is not None:
        args += ["--preprocessed", str(preprocessed)]
    return args


@pytest.fixture()
def files(tmp_path, brp):
    n = 1 << 16
    block, _ = brp.synthetic_dd(n, 64e-6, 21, 7.5, 2.0, 4, 200.0, 0.05, 1.0, (3600.0, 0.05, 1.0), n, 4, 50.0)
    dd = write_dd(tmp_path / "obs.bin4", block, n)
    bank = tmp_path / "bank.txt"
    rows = [(3600.0, 0.05, 1.0)] + [(o.period, o.tau, o.psi) for o in brp.synthetic_bank(5, 23)]
    bank.write_text("".join(f"{p!r} {t!r} {s!r}\n" for p, t, s in rows))
    zap = tmp_path / "zap.txt"
    zap.write_text("60.0 60.5\n1000 1001.25\n")
    return {"n": n, "block": block, "dd": dd, "bank": bank, "zap": zap, "rows": rows}


# ------------------------------------------------------------------ CPU only --

def test_bench_help_and_argument_errors():
    r = run([exe("brp_bench"), "--help"])
    assert r.returncode == 0 and r.stdout.startswith("brp_bench [--n 2^k|3*2^k]")
    for args, msg in ((["--bogus", "1"], "unknown option: --bogus"), (["--reps", "0"], "invalid repetitions"),
                      (["--n"], "option needs value"), (["--engine", "cuda"], "--engine must be fused or plain"),
                      (["--n", "100"], "invalid observation/template")):
        r = run([exe("brp_bench"), *args])
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)


def test_workunit_argument_errors(tmp_path, files):
    base = workunit_args(tmp_path, files["dd"], files["bank"], files["zap"])
    cases = [
        (base[:3], "required option --bank"),
        (base + ["--bogus", "1"], "unknown option --bogus"),
        (base + ["--wg", "64", "--wg", "64"], "duplicate option"),
        (base + ["--cpu"], "test infrastructure"),
        (base + ["--engine", "vulkan"], "--engine must be fused or plain"),
        (base + ["--radix8", "--dit"], "radix 8 requires Stockham"),
        (base + ["--validate", "--batch", "4"], "intermediate validation requires batch 1"),
        (base + ["--batch"], "option needs a value"),
    ]
    for args, msg in cases:
        r = run(args)
        assert r.returncode == 1 and msg in r.stderr, (args[-3:], r.stderr)
    (tmp_path / "out.txt").write_text("x")
    r = run(base)
    assert r.returncode == 1 and "refusing to overwrite results" in r.stderr


def test_workunit_dd_format_errors(tmp_path, files):
    """The DD loader's checks (src/workload.cpp:29-56), all before any device call."""
    n, block = files["n"], files["block"]
    cases = {
        "trunc.bin4": (lambda p: p.write_bytes(gzip.compress(b"\0" * 100)), "truncated DD header"),
        "short.bin4": (lambda p: write_dd(p, block[:100], n), "truncated DD samples"),
        "extra.bin4": (lambda p: write_dd(p, block, n, extra=b"\1"), "trailing DD bytes"),
        "odd.bin4": (lambda p: write_dd(p, block, n - 1), "odd packed DD length"),
        "obs.dat": (lambda p: write_dd(p, block, n), "unsupported DD sample format"),
        "scale.bin4": (lambda p: write_dd(p, block, n, scale=0.0), "invalid DD metadata"),
        "big.bin4": (lambda p: write_dd(p, block, 3 * (1 << 22) + 2), "observation exceeds supported length"),
        "missing.bin4": (lambda p: None, "cannot open DD input"),
    }
    for name, (make, msg) in cases.items():
        path = tmp_path / name
        make(path)
        r = run(workunit_args(tmp_path, path, files["bank"], files["zap"]))
        assert r.returncode == 1 and msg in r.stderr, (name, r.stderr)


def test_workunit_bank_errors_and_uncompressed_input(tmp_path, files):
    """Bank rows (src/workload.cpp:96-107) are checked before the device is used when the
    preprocessed series is given; the DD file may also be uncompressed."""
    n = files["n"]
    pre = tmp_path / "pre.f32"
    np.zeros(n, dtype=np.float32).tofile(pre)
    plain_dd = write_dd(tmp_path / "plain.bin4", files["block"], n, gz=False)
    for text, msg in (("", "empty orbital bank"), ("100 0.1\n", "invalid orbital bank row"),
                      ("100 0.1 1 2\n", "invalid orbital bank row"), ("-5 0.1 1\n", "invalid orbital bank row"),
                      ("100 -0.1 1\n", "invalid orbital bank row"), ("100 nan 1\n", "invalid orbital bank row")):
        bank = tmp_path / "bad_bank.txt"
        bank.write_text(text)
        r = run(workunit_args(tmp_path, plain_dd, bank, files["zap"], preprocessed=pre))
        assert r.returncode == 1 and msg in r.stderr, (text, r.stderr)
    short = tmp_path / "short.f32"
    np.zeros(n - 1, dtype=np.float32).tofile(short)
    r = run(workunit_args(tmp_path, plain_dd, files["bank"], files["zap"], preprocessed=short))
    assert r.returncode == 1 and "reference float dimensions" in r.stderr


# ----------------------------------------------------------------------- GPU --

def _toplist_file(brp, files, tmp_path, rfi_ratio, engine="fused"):
    """The same workunit through the library: device preprocessing with the zap list,
    brp_hip_bank_search in batches of 64 (brp_workunit's defaults: workgroup 256, Stockham,
    radix 4), final rows written by brp_write_candidates."""
    n = files["n"]
    zap = [tuple(map(float, line.split())) for line in files["zap"].read_text().splitlines()]
    x, _, rfi = brp.preprocess(files["block"], n, 4, 1.0, n, 1000, rfi_ratio, zap, 64.0, 1.0)
    orbits = [brp.Orbit(*r) for r in files["rows"]]
    first = brp.search_parameters(n, n, 64.0, orbits[0])
    top = brp.Toplist(100)
    kw = dict(engine=brp.ENGINE_FUSED, graph=True) if engine == "fused" else dict(engine=brp.ENGINE_PLAIN)
    with brp.GPU(first, x, 256, 0, 64, stockham=True, radix4=True, packed=True, fused_power=True, **kw) as g:
        g.bank_search(orbits, n, 64.0, top)
    path = tmp_path / "lib.txt"
    brp.write_candidates(str(path), top.final(), n * 64.0 * 1e-6)
    return path.read_text(), rfi


@pytest.mark.gpu
def test_workunit_end_to_end_matches_library(brp, tmp_path, files):
    """brp_workunit on a 4-bit DD file with a zap list and RFI replacement: exit 0, one
    metrics object per template, and a results file identical to the library pipeline's."""
    say("[gpu] brp_workunit N=2^16, 24 templates, zap list, rfi 30")
    r = run(workunit_args(tmp_path, files["dd"], files["bank"], files["zap"], ["--rfi-ratio", "30"]))
    say(r.stdout, r.stderr[-500:])
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout)
    assert summary["templates"] == 24 and summary["valid"] and summary["rfi_bins"] >= 4
    metrics = [json.loads(line) for line in (tmp_path / "metrics.jsonl").read_text().splitlines()]
    assert [m["template"] for m in metrics] == list(range(24))
    assert all(len(m["thresholds"]) == 5 and m["n"] == files["n"] for m in metrics)
    text = (tmp_path / "out.txt").read_text()
    assert text.startswith("% Experimental offline BRP4 results") and text.endswith("%DONE%\n")
    expected, rfi = _toplist_file(brp, files, tmp_path, 30.0)
    assert rfi == summary["rfi_bins"]
    assert text == expected
    assert summary["candidates"] == len(text.splitlines()) - 2 > 0


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ["fused", "plain"])
def test_workunit_validate_cross_checks_engines(brp, tmp_path, files, engine):
    """--validate (batch 1): each template checked against the other engine; all valid,
    and the results file equals the same engine's batched run (frozen thresholds give
    the one-at-a-time toplist).  Across engines the last-ulp power differences can
    reorder rows that tie at the saturated significance, so files are compared within
    one engine."""
    say(f"[gpu] brp_workunit --validate --engine {engine}")
    r = run(workunit_args(tmp_path, files["dd"], files["bank"], files["zap"],
                          ["--validate", "--batch", "1", "--engine", engine, "--packed", "--fused-power"]))
    assert r.returncode == 0, r.stderr
    metrics = [json.loads(line) for line in (tmp_path / "metrics.jsonl").read_text().splitlines()]
    assert all(m["valid"] and m["validated"] and m["sampled_mismatches"] == 0 for m in metrics)
    assert max(m["power_l2"] for m in metrics) < 5e-5
    expected, _ = _toplist_file(brp, files, tmp_path, 0.0, engine)
    assert (tmp_path / "out.txt").read_text() == expected


@pytest.mark.gpu
@pytest.mark.parametrize("args", [["--n", "65536", "--reps", "3"],
                                  ["--n", "98304", "--reps", "3", "--production", "--batch", "8"],
                                  ["--n", "65536", "--reps", "3", "--engine", "plain", "--stockham", "--tile", "0",
                                   "--radix8", "--packed", "--fused-power"]])
def test_bench_driver(args):
    """brp_bench: exit 0 with valid true, the reference's JSON fields, and the engines
    agreeing bit for bit on the resample indices."""
    say(f"[gpu] brp_bench {' '.join(args)}")
    r = run([exe("brp_bench"), *args])
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    for key in ("device", "n", "steps", "valid", "index_mismatches", "candidate_mismatches", "rank_mismatches",
                "replay_reproducible", "resample_l2", "fft_l2", "power_l2", "harmonic_l2", "wall_ms_median",
                "gpu_ms_median", "replays_per_second", "stage_ms_median", "wall_ms_samples", "hardware_counters"):
        assert key in out, key
    assert out["valid"] and out["replay_reproducible"] and out["index_mismatches"] == 0
    assert len(out["wall_ms_samples"]) == 3 and all(math.isfinite(v) and v > 0 for v in out["wall_ms_samples"])


@pytest.mark.gpu
def test_workunit_full_size(brp, tmp_path):
    """brp_workunit at the production size: a 4-bit DD file of 2^22 samples and a bank of
    512 orbits; the results file equals the library pipeline's byte for byte."""
    n = 1 << 22
    block, _ = brp.synthetic_dd(n, 64e-6, 42, 7.5, 2.0, 4, 211.0, 0.05, 0.2, (3600.0, 0.05, 1.0), n, 10, 50.0)
    dd = write_dd(tmp_path / "obs.bin4", block, n)
    rows = [(3600.0, 0.05, 1.0)] + [(o.period, o.tau, o.psi) for o in brp.synthetic_bank(43, 511)]
    bank = tmp_path / "bank.txt"
    bank.write_text("".join(f"{p!r} {t!r} {s!r}\n" for p, t, s in rows))
    zap = tmp_path / "zap.txt"
    zap.write_text("50.0 50.2\n")
    files = {"n": n, "block": block, "dd": dd, "bank": bank, "zap": zap, "rows": rows}
    say("[gpu] brp_workunit N=2^22, 512 templates")
    r = run(workunit_args(tmp_path, dd, bank, zap, ["--rfi-ratio", "30", "--batch", "128"]), timeout=300)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout)
    say(summary)
    assert summary["templates"] == 512 and summary["valid"] and summary["candidates"] > 0
    expected, rfi = _toplist_file(brp, files, tmp_path, 30.0)
    assert rfi == summary["rfi_bins"]
    assert (tmp_path / "out.txt").read_text() == expected
