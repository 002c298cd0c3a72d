"""Host-side science of the product (csrc/science.cpp) against the oracle, on the CPU."""
import ctypes
import json
import math
import pathlib
import re

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_product_lut_matches_fixture():
    fx = json.loads((ROOT / "tests" / "golden" / "lut_table.json").read_text())
    text = (ROOT / "boinc-einstein-bpr4-v3d_amd" / "csrc" / "lut.h").read_text()
    body = text[text.index("kTable[130] = {") + 15:text.index("};")]
    vals = np.float32([float(v.strip().rstrip("f")) for v in body.split(",") if v.strip()])
    assert np.array_equal(vals[0::2], np.float32(fx["sinSamples"]))
    assert np.array_equal(vals[1::2], np.float32(fx["cosSamples"]))
    assert np.array_equal(np.signbit(vals[0::2]), np.signbit(np.float32(fx["sinSamples"])))


def test_lut_sin_and_offset_equal_oracle(brp, oracle):
    lib = oracle.load()
    xs = np.random.default_rng(11).uniform(-30, 30, 5000).astype(np.float32)
    assert all(brp.lut_sin(float(x)) == lib.oracle_lut_sin(float(x)) for x in xs)
    p = brp.search_parameters(1 << 20, 1 << 20, 64.0, (700.0, 0.28, 2.2))
    op = oracle.as_oparams(p)
    for i in range(0, 1 << 20, 4099):
        assert brp.sample_offset(i, p) == lib.oracle_sample_offset(i, op)


def _bytes(p):
    return bytes(ctypes.string_at(ctypes.addressof(p), 108))


@pytest.mark.parametrize("n", [1 << 16, 3 << 15, 1 << 22])
def test_parameters_equal_oracle(brp, oracle, n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        tau, period, psi = rng.uniform(0, 0.3), math.exp(rng.uniform(math.log(660), math.log(10000))), rng.uniform(0, 6.28)
        a = brp.parameters(n, n - 17, tau, period, psi, 64e-6)
        b = oracle.parameters(n, n - 17, tau, period, psi, 64e-6)
        assert _bytes(a) == _bytes(b)
        a = brp.search_parameters(n, n, 64.0, (period, tau, psi))
        b = oracle.search_parameters(n, n, 64.0, (period, tau, psi))
        assert _bytes(a) == _bytes(b)


def test_resampling_steps_equal_oracle_on_extreme_orbits(brp, oracle):
    # Short periods and large tau (tau * omega up to ~0.19) stress the jump-started scan.
    rng = np.random.default_rng(99)
    for n in (1 << 14, 3 << 13, 1 << 20):
        for _ in range(40):
            tau, period, psi = rng.uniform(0, 0.3), rng.uniform(10, 200), rng.uniform(0, 6.28)
            a = brp.parameters(n, n - int(rng.integers(0, 1000)), tau, period, psi, 64e-6)
            b = oracle.parameters(n, a.unpadded, tau, period, psi, 64e-6)
            assert _bytes(a) == _bytes(b)


def test_resampling_steps_long_delays_equal_the_linear_scan(brp, oracle):
    """Delays of 10^4 .. 2 x 10^7 samples (tau up to 1280 s at dt = 64 us, tau omega < 0.5):
    past the jump start's error bound the reference's scan from unpadded - 1 runs, and every
    result equals the oracle's linear scan (src/science.cpp:41-45)."""
    rng = np.random.default_rng(7)
    n = 1 << 22
    for tau in (0.64, 6.4, 32.0, 64.0, 160.0, 320.0):
        for _ in range(3):
            period = float(rng.uniform(2 * np.pi * tau / 0.45, 4 * np.pi * tau / 0.45))
            psi = float(rng.uniform(0, 6.28))
            a = brp.parameters(n, n - int(rng.integers(0, 1000)), tau, period, psi, 64e-6)
            b = oracle.parameters(n, a.unpadded, tau, period, psi, 64e-6)
            assert _bytes(a) == _bytes(b), (tau, period, psi)
    # Orbits whose FP32 rule lies 30-140 samples beyond the FP64 solution: a scan started
    # 64 samples above that solution would stop short of the reference's length.
    for tau, period, psi in ((640.0, 10219.499304313382, 2.7828623112678965),
                             (640.0, 14349.413763432298, 3.783723419434997),
                             (1280.0, 28940.997487526118, 2.8567255598987424)):
        a = brp.parameters(n, n, tau, period, psi, 64e-6)
        b = oracle.parameters(n, n, tau, period, psi, 64e-6)
        assert _bytes(a) == _bytes(b), (tau, period, psi)


def test_parameter_validation_messages(brp):
    with pytest.raises(RuntimeError, match="invalid observation/template"):
        brp.parameters(100, 100, 0.01, 100.0, 0.1, 64e-6)
    with pytest.raises(RuntimeError, match="invalid search parameters"):
        brp.search_parameters(1 << 16, 1 << 16, 64.0, (100.0, 0.01, 0.1), false_alarm=1.5)


@pytest.mark.parametrize("ties", [False, True])
def test_toplist_equals_oracle_on_random_streams(brp, oracle, ties):
    """The incremental stable insertion equals the reference's rewrite + stable sort, also
    with many equal powers (ties keep their order of arrival)."""
    rng = np.random.default_rng(12 + ties)
    n = 1 << 20
    for rows in (3, 7, 100):
        top, otop = brp.Toplist(rows), oracle.OToplist(rows)
        for t in range(60):
            orbit = (float(rng.uniform(660, 10000)), float(rng.uniform(0, 0.3)), float(rng.uniform(0, 6.28)))
            p = brp.search_parameters(n, n, 64.0, orbit)
            op = oracle.search_parameters(n, n, 64.0, orbit)
            top.thresholds(p)
            otop.thresholds(op)
            assert list(p.threshold) == list(op.threshold)
            k = int(rng.integers(0, 12))
            levels = rng.integers(0, 5, k)
            bins = rng.integers(p.window, p.fundamental, k)
            bins[: k // 3] = rng.integers(p.window, p.window + 5, k // 3)  # repeated bins
            if ties:  # a few distinct levels of power, far above every threshold
                powers = np.float32([100.0 * int(rng.integers(1, 4)) for _ in levels])
            else:
                powers = np.float32([p.threshold[l] * rng.uniform(0.9, 3.0) for l in levels])
            cands = np.zeros(k, dtype=brp.CANDIDATE_DTYPE)
            cands["bin"], cands["level"], cands["power"] = bins, levels, powers
            cands = np.sort(cands, order=["level", "bin"])
            top.retain([brp.Candidate(int(c["bin"]), int(c["level"]), float(c["power"]), 0) for c in cands], orbit, p)
            otop.retain(cands, orbit, op)
        mine = [(r.bin, r.power, r.period, r.tau, r.psi, r.harmonics) for r in top.table()]
        ref = [(r.bin, r.power, r.period, r.tau, r.psi, r.harmonics) for r in otop.state]
        assert mine == ref
        fa, fb = top.final(), otop.final()
        assert [(r.bin, r.power, r.harmonics) for r in fa] == [(r.bin, r.power, r.harmonics) for r in fb]
        # Independent log-domain evaluations of the chi^2 tail: equal to a few ulp.
        assert np.allclose([r.significance for r in fa], [r.significance for r in fb], rtol=1e-12)
        assert len(fa) > 0


def test_write_candidates_format(brp, tmp_path):
    rows = [brp.BankRow(12.5, 3600.0, 0.05, 1.0, 9.25, 4, 4242)]
    path = tmp_path / "results.cand"
    brp.write_candidates(str(path), rows, 268.435456)
    lines = path.read_text().splitlines()
    assert lines[0].startswith("%") and lines[-1] == "%DONE%"
    f = lines[1].split()
    assert float(f[0]) == pytest.approx(4242 / 268.435456) and int(f[6]) == 4 and float(f[4]) == 12.5


def test_synthetic_series_statistics_and_pulsar(brp):
    n, dt = 1 << 18, 64e-6
    a = brp.synthetic_series(n, dt, 42)
    b = brp.synthetic_series(n, dt, 42)
    c = brp.synthetic_series(n, dt, 43)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1) < 0.01
    assert abs(np.mean(a ** 4) / np.mean(a ** 2) ** 2 - 3) < 0.05  # Gaussian kurtosis
    # A strong un-Doppler-shifted train concentrates power at its harmonics.
    x = brp.synthetic_series(n, dt, 7, spin_hz=200.0, duty=0.05, amplitude=1.0, orbit=(3600.0, 0.0, 0.0))
    spec = np.abs(np.fft.rfft(x.astype(np.float64))) ** 2 / n
    f_bin = 200.0 * n * dt
    k = int(round(f_bin))
    assert spec[k - 1:k + 2].max() > 50 * np.median(spec)
    # The orbit model: sample j of the train sits at pulsar time t_p with
    # t_p - tau sin(w t_p + psi) + tau sin psi = j dt; a matching resample makes it periodic.
    period, tau, psi = 600.0, 0.2, 1.1
    x = brp.synthetic_series(n, dt, 8, sigma=1.0, mean=0.0, spin_hz=100.0, duty=0.25, amplitude=100.0,
                             orbit=(period, tau, psi))
    j = np.arange(n)
    w = 2 * np.pi / period
    tp = j * dt
    for _ in range(8):
        tp = j * dt + tau * (np.sin(w * tp + psi) - np.sin(psi))
    on = (100.0 * tp) % 1.0 < 0.25
    assert np.array_equal(x > 50.0, on)


def test_synthetic_bank_ranges(brp):
    bank = brp.synthetic_bank(5, 2000)
    per = np.array([o.period for o in bank]); tau = np.array([o.tau for o in bank]); psi = np.array([o.psi for o in bank])
    assert per.min() >= 660 and per.max() <= 10000 and tau.min() >= 0 and tau.max() <= 0.3
    assert psi.min() >= 0 and psi.max() < 2 * np.pi
    # log-uniform period: the median sits near the geometric mean.
    assert abs(np.log(np.median(per)) - 0.5 * (np.log(660) + np.log(10000))) < 0.15
