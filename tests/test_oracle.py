"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

The reference's own CPU path cannot be built here (FFTW3f, GSL and the historical
boinc-app-eah-brp sources are absent), and the reference ships no golden vectors for
this path, so each oracle function is pinned by checks that a plausible mistake
(dropped term, wrong sign or index, transposed operand) would fail:

  lut table        fixture parsed from the reference header (tests/golden/lut_table.json)
  lut_sin          |lut_sin - sin| <= Taylor remainder + table rounding; mismatch of a term fails
  sample_offset    vs the exact FP64 orbit model, within the LUT error bound
  resample         tau = 0 is the identity; |i - j - off_exact| <= 1/2 + bound; y = x[j]
  resampling_steps brute force over every candidate length
  FFT / power      numpy.fft.rfft (FP64 pocketfft), impulse, integer-bin cosine, Parseval,
                   brute-force DFT for 3*2^k
  harmonic         integer-valued spectra (exact in FP32) against the textbook sum
                   sum_m P[round(i m / 2^h)] with the max over each rounding interval;
                   gather == scatter bit-exact on random spectra
  chi^2            scipy.stats.chi2.sf / isf
  toplist          hand-built replacement / ordering / cross-level cases
  preprocess       independent numpy restatement (numpy.fft.rfft / irfft, numpy.median of
                   each window), decode bit for bit, injected RFI bins found and replaced
"""
import json
import math
import pathlib

import numpy as np
import pytest
from scipy import stats

GOLDEN = pathlib.Path(__file__).with_name("golden")


def test_lut_table_matches_fixture(oracle):
    fx = json.loads((GOLDEN / "lut_table.json").read_text())
    t = oracle.lut_table()
    assert np.array_equal(t[0::2], np.float32(fx["sinSamples"]))
    assert np.array_equal(t[1::2], np.float32(fx["cosSamples"]))
    assert np.array_equal(np.signbit(t[0::2]), np.signbit(np.float32(fx["sinSamples"])))


def test_lut_sin_error_bound(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-40, 40, 20000), np.linspace(0, 2 * math.pi, 4097)]).astype(np.float32)
    lib = oracle.load()
    got = np.array([lib.oracle_lut_sin(float(x)) for x in xs])
    fgot = np.array([lib.oracle_fused_sin(float(x)) for x in xs])
    exact = np.sin(xs.astype(np.float64))
    # |d| <= pi/64: Taylor remainder |d|^3/6 ~ 2.0e-5, table entries rounded to 1e-6.
    bound = (math.pi / 64) ** 3 / 6 + 2.5e-6
    assert np.max(np.abs(got - exact)) < bound
    assert np.max(np.abs(fgot - exact)) < bound
    # The second-order term matters: without it the error would reach (pi/64)^2/2.
    assert np.max(np.abs(got - exact)) < 0.02 * (math.pi / 64) ** 2 / 2 * 10
    # Periodicity (modf + wrap): x and x - 2 pi agree to FP32 rounding of the argument.
    for x in (0.3, 1.7, 3.1, 5.9):
        assert abs(lib.oracle_lut_sin(x) - lib.oracle_lut_sin(x - 2 * math.pi)) < 1e-5


def _params(oracle, n, unpadded, tau, period, psi, dt, flags=0):
    p = oracle.parameters(n, unpadded, tau, period, psi, dt)
    if flags:
        p.flags = flags
        p.s0 = np.float32(np.float32(p.tau) * np.sin(np.float32(p.psi))) * np.float32(p.step_inv)
        oracle.load().oracle_resampling_steps(p)
    return p


@pytest.mark.parametrize("flags", [0, 6])
def test_sample_offset_tracks_exact_orbit(oracle, flags):
    tau, period, psi, dt = 0.3, 700.0, 2.5, 64e-6
    p = _params(oracle, 1 << 20, 1 << 20, tau, period, psi, dt, flags)
    lib = oracle.load()
    idx = np.arange(0, 1 << 20, 997, dtype=np.uint32)
    got = np.array([lib.oracle_sample_offset(int(i), p) for i in idx], dtype=np.float64)
    t = idx.astype(np.float64) * np.float32(dt)
    exact = tau / dt * np.sin(2 * np.pi / period * t + psi) - float(p.s0)
    bound = (tau / dt) * ((math.pi / 64) ** 3 / 6 + 3e-6) + 0.01
    assert np.max(np.abs(got - exact)) < bound
    # A sign error in either orbit term moves the offset by ~2 tau/dt.
    assert bound < 0.1 * tau / dt


def test_resample_identity_for_zero_tau(oracle):
    x = np.random.default_rng(2).standard_normal(4096).astype(np.float32)
    p = oracle.parameters(4096, 4096, 0.0, 1000.0, 0.7, 64e-6)
    # Zero offset: the largest length with float(s) - 0 < U - 1 is U - 2 (strict test).
    assert p.steps == 4094
    y, j = oracle.resample(x, p)
    assert np.array_equal(j, np.arange(4094))
    assert np.array_equal(y[:4094], x[:4094])
    acc = np.float32(0)
    for v in x[:4094]:
        acc = np.float32(acc + v)
    assert np.all(y[4094:] == np.float32(acc / np.float32(4094)))


@pytest.mark.parametrize("case", [(0.05, 3600.0, 1.0), (0.3, 660.0, 4.0), (0.2, 1500.0, 5.5)])
def test_resample_indices_round_the_exact_model(oracle, case):
    tau, period, psi = case
    dt, n = 64e-6, 1 << 18
    x = np.random.default_rng(3).standard_normal(n).astype(np.float32)
    p = oracle.parameters(n, n, tau, period, psi, dt)
    y, j = oracle.resample(x, p)
    i = np.arange(p.steps, dtype=np.float64)
    t = i * np.float32(dt)
    off = tau / dt * np.sin(2 * np.pi / period * t + psi) - float(p.s0)
    resid = i - off - j
    bound = (tau / dt) * ((math.pi / 64) ** 3 / 6 + 3e-6) + 0.02
    assert np.all(np.abs(resid) <= 0.5 + bound)
    near = np.abs(np.abs(resid) - 0.5) < bound + 1e-9
    exact = np.floor(i - off + 0.5)
    assert np.all((exact == j) | near)
    assert np.array_equal(y[:p.steps], x[j])
    # The pad is the FP32 mean of the resampled samples.
    assert np.all(y[p.steps:] == y[p.steps])
    assert abs(float(y[p.steps]) - float(np.mean(y[:p.steps], dtype=np.float64))) < 1e-5


def test_resampling_steps_is_the_largest_valid_length(oracle):
    lib = oracle.load()
    for tau, period, psi in [(0.05, 3600.0, 1.0), (0.3, 700.0, 2.0), (0.25, 900.0, 4.5)]:
        n = 1 << 16
        p = oracle.parameters(n, n, tau, period, psi, 64e-6)
        last = np.float32(n - 1)
        ok = lambda s: np.float32(s) - np.float32(lib.oracle_sample_offset(s, p)) < last
        assert ok(p.steps)
        assert not any(ok(s) for s in range(p.steps + 1, n))


def test_parameters_fields_and_validation(oracle):
    p = oracle.parameters(3 * (1 << 15), 3 * (1 << 15) - 17, 0.002, 120.0, 0.3, 0.000256)
    assert (p.fft_radix, p.fft_base, p.logn) == (3, 1 << 15, 15)
    assert p.omega == np.float32(2 * math.pi / np.float32(120.0))
    assert p.s0 == np.float32(np.float64(np.float32(0.002)) * math.sin(np.float64(np.float32(0.3))) *
                              np.float64(np.float32(1) / np.float32(0.000256)))
    assert (p.fundamental, p.harmonic, p.window, p.capacity) == (3 << 14, 3 << 14, 8, 65536)
    assert list(p.threshold) == [20, 30, 45, 70, 110]
    lib = oracle.load()
    bad = oracle.OParams()
    for args in [(100, 100), (5 * 1024, 5 * 1024), (32, 32), (1 << 16, 16), (1 << 16, (1 << 16) + 1)]:
        assert lib.oracle_parameters(args[0], args[1], 0.01, 100.0, 0.1, 64e-6, bad) != 0


@pytest.mark.parametrize("n", [64, 96, 1 << 12, 3 << 10, 1 << 16, 3 << 14])
def test_rfft64_matches_numpy(oracle, n):
    x = np.random.default_rng(n).standard_normal(n)
    got = oracle.rfft64(x)
    ref = np.fft.rfft(x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.sqrt(n) * np.max(np.abs(ref))


def test_rfft64_closed_forms(oracle):
    n = 3 << 8
    imp = np.zeros(n); imp[5] = 1.0
    got = oracle.rfft64(imp)
    k = np.arange(n // 2 + 1)
    assert np.allclose(got, np.exp(-2j * np.pi * 5 * k / n), atol=1e-13)
    cos = np.cos(2 * np.pi * 37 * np.arange(n) / n)
    spec = oracle.rfft64(cos)
    assert abs(spec[37] - n / 2) < 1e-9 and np.max(np.abs(np.delete(spec, 37))) < 1e-9
    x = np.random.default_rng(5).standard_normal(n)
    s = oracle.rfft64(x)
    parseval = (abs(s[0]) ** 2 + abs(s[-1]) ** 2 + 2 * np.sum(np.abs(s[1:-1]) ** 2)) / n
    assert abs(parseval - np.sum(x * x)) < 1e-9 * np.sum(x * x)
    # Brute-force DFT for a 3 * 2^k length.
    n = 3 * 32
    x = np.random.default_rng(6).standard_normal(n)
    dft = np.exp(-2j * np.pi * np.outer(np.arange(n // 2 + 1), np.arange(n)) / n) @ x
    assert np.allclose(oracle.rfft64(x), dft, atol=1e-11)


def test_rfft32_is_accurate(oracle):
    for n in (1 << 16, 3 << 14):
        x = np.random.default_rng(7).standard_normal(n).astype(np.float32)
        got = oracle.rfft32(x).astype(np.complex128)
        ref = np.fft.rfft(x.astype(np.float64))
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-6


def test_power_definition(oracle):
    n = 1 << 12
    x = np.random.default_rng(8).standard_normal(n).astype(np.float32)
    p32, p64 = oracle.power(x)
    ref = np.abs(np.fft.rfft(x.astype(np.float64))[: n // 2]) ** 2 / n
    ref[0] = 0
    assert p32[0] == 0 and p64[0] == 0
    assert np.allclose(p64, ref, rtol=1e-12, atol=1e-12)
    assert np.array_equal(p32, p64.astype(np.float32))


def _textbook_harmonic(P, F, H, W, thr):
    """sum_m P[round_half_up(i m / 2^h)], max over the rounding interval of j, first-value rule."""
    out = np.zeros((5, F), np.float64)
    out[0] = P[:F]
    for h in range(1, 5):
        width = 1 << h
        for j in range(F):
            begin, end = max(W, j * width - width // 2), min(H, j * width + width // 2)
            if begin >= end:
                continue
            vals = [sum(P[(i * m * (16 // width) + 8) >> 4] for m in range(1, width + 1)) for i in range(begin, end)]
            best = max(vals)
            out[h, j] = best if best > thr[h] else vals[0]
    return out


def test_harmonic_against_textbook_definition(oracle):
    rng = np.random.default_rng(9)
    n = 2048
    P = rng.integers(0, 9, n // 2).astype(np.float32)  # small integers: every FP32 sum is exact
    P[137] = 60; P[274] = 45; P[411] = 40  # a harmonic family
    p = oracle.parameters(n, n, 0.0, 100.0, 0.0, 64e-6)
    p.fundamental, p.harmonic, p.window = 300, 1000, 8
    thr = [20.0, 30.0, 45.0, 70.0, 110.0]
    for h in range(5):
        p.threshold[h] = thr[h]
    got = oracle.harmonic(P, p).reshape(5, p.fundamental)
    ref = _textbook_harmonic(P.astype(np.int64), 300, 1000, 8, thr)
    assert np.array_equal(got, ref.astype(np.float32))
    cands = oracle.select(got.ravel(), p)
    ref_c = [(j, h) for h in range(5) for j in range(8, 300) if ref[h, j] > thr[h]]
    assert [(int(c["bin"]), int(c["level"])) for c in cands] == ref_c
    assert len(ref_c) > 3


def test_harmonic_gather_equals_scatter(oracle):
    rng = np.random.default_rng(10)
    for n, F, H, W in [(1 << 14, 1 << 13, 1 << 13, 8), (1 << 14, 500, 7000, 50)]:
        P = rng.exponential(1.0, n // 2).astype(np.float32)
        P[0] = 0
        p = oracle.parameters(n, n, 0.0, 100.0, 0.0, 64e-6)
        p.fundamental, p.harmonic, p.window = F, H, W
        for h, t in enumerate([6.0, 8.0, 11.0, 16.0, 25.0]):
            p.threshold[h] = t
        a = oracle.harmonic(P, p)
        b = oracle.harmonic(P, p, scatter=True)
        assert np.array_equal(a, b)
        assert len(oracle.select(a, p)) > 10


def test_chisq_closed_form_against_scipy(oracle):
    for dof in (2, 4, 8, 16, 32):
        for x in (0.5, 3.0, 20.0, 80.0, 300.0):
            q = oracle.chisq_Q(x, dof)
            ref = stats.chi2.sf(x, dof)
            assert abs(q - ref) <= 1e-10 * ref + 1e-300
        for q in (0.5, 1e-3, 3.976e-8, 1e-12):
            assert abs(oracle.chisq_Qinv(q, dof) - stats.chi2.isf(q, dof)) < 1e-8 * stats.chi2.isf(q, dof)


def test_search_parameters_bank_semantics(oracle):
    # Values computed independently with scipy.stats.chi2.isf for f0 = 400 Hz, false alarm 0.08.
    table = {
        1 << 20: (500, 26844, 429497, [15.654, 18.631, 23.458, 31.525, 45.362]),
        3 << 19: (500, 40265, 644245, [16.060, 19.058, 23.920, 32.040, 45.957]),
        3 << 20: (500, 80531, 1288490, [16.753, 19.787, 24.706, 32.915, 46.966]),
        1 << 22: (500, 107374, 1717987, [17.040, 20.089, 25.031, 33.276, 47.382]),
    }
    for n, (W, F, H, thr) in table.items():
        p = oracle.search_parameters(n, n, 64.0, (3600.0, 0.05, 1.0))
        assert (p.window, p.fundamental, p.harmonic, p.flags) == (W, F, H, 6)
        assert np.allclose(list(p.threshold), thr, atol=1.5e-3)
        prob = np.float32(1.0 - (1.0 - 0.08) ** (1.0 / (n // 2 + 1)))
        for h in range(5):
            assert p.threshold[h] == np.float32(0.5 * stats.chi2.isf(float(prob), 2 * 2 ** h)) or \
                abs(p.threshold[h] - 0.5 * stats.chi2.isf(float(prob), 2 * 2 ** h)) < 4e-6 * p.threshold[h]


def _cand(bin_, level, power):
    return np.array([(bin_, level, power, 0)], dtype=oracle_dtype())[0]


def oracle_dtype():
    return np.dtype([("bin", "<u4"), ("level", "<u4"), ("power", "<f4"), ("reserved", "<u4")])


def test_toplist_replacement_and_final(oracle):
    p = oracle.search_parameters(1 << 20, 1 << 20, 64.0, (3600.0, 0.05, 1.0))
    top = oracle.OToplist(rows=3)
    c = lambda rows: np.array(rows, dtype=oracle_dtype())
    top.retain(c([(1000, 0, 20.0, 0), (1001, 0, 30.0, 0), (1002, 0, 25.0, 0)]), (100.0, 0.1, 0.2), p)
    # Same bin, weaker: ignored.  Same bin, stronger: replaced in place.  New bin: evicts the last.
    top.retain(c([(1000, 0, 19.0, 0)]), (200.0, 0.1, 0.2), p)
    top.retain(c([(1002, 0, 40.0, 0)]), (300.0, 0.1, 0.2), p)
    top.retain(c([(1003, 0, 21.0, 0)]), (400.0, 0.1, 0.2), p)
    rows = [(r.bin, r.power, r.period) for r in top.state[:3]]
    assert rows == [(1002, 40.0, 300.0), (1001, 30.0, 100.0), (1003, 21.0, 400.0)]
    # Below threshold is dropped even on an empty level.
    top.retain(c([(1004, 1, 10.0, 0)]), (500.0, 0.1, 0.2), p)
    assert top.state[3].harmonics == 0
    # Level 1 holds bin 1002 too: final keeps the most significant one per bin.
    top.retain(c([(1002, 1, 60.0, 0)]), (600.0, 0.1, 0.2), p)
    fin = top.final()
    bins = [r.bin for r in fin]
    assert len(bins) == len(set(bins)) == 3
    sig = {r.bin: r.significance for r in fin}
    assert sig[1002] == pytest.approx(-math.log10(stats.chi2.sf(120.0, 4)), rel=1e-9)
    f1002 = [r for r in fin if r.bin == 1002][0]
    assert f1002.harmonics == 2 and f1002.power == pytest.approx(60.0 / math.sqrt(2.0), rel=1e-6)
    # Thresholds rise to the last row of each level.
    q = oracle.search_parameters(1 << 20, 1 << 20, 64.0, (3600.0, 0.05, 1.0))
    top.thresholds(q)
    assert q.threshold[0] == np.float32(21.0)


_MASK64 = (1 << 64) - 1


def _splitmix_mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return z ^ (z >> 31)


def _uniform(seed, counter):
    """SplitMix64 counter-based uniform, written from the published finaliser."""
    z = _splitmix_mix((_splitmix_mix((seed + 0x9E3779B97F4A7C15) & _MASK64) +
                       counter * 0xD1B54A32D192ED03) & _MASK64)
    return (z >> 11) * 2.0 ** -53


def _numpy_preprocess(block, samples, bits, scale, n, window, rfi_ratio, zap=(), sample_us=0.0, padding=1.0):
    if bits == 4:
        v = np.empty(samples, dtype=np.float64)
        v[0::2] = block >> 4
        v[1::2] = block & 15
    else:
        v = block.view(np.int8).astype(np.float64)
    dec = (v / scale).astype(np.float32)
    X = np.fft.rfft(np.concatenate([dec.astype(np.float64), np.zeros(n - samples)]))
    P = (X.real ** 2 + X.imag ** 2).astype(np.float32)
    P[0] = 0
    win = np.lib.stride_tricks.sliding_window_view(P.astype(np.float64), window)
    med = np.median(win, axis=1).astype(np.float32)
    Y = np.zeros_like(X)
    rfi = 0
    for k in range(window // 2, len(X) - window // 2):
        m = med[k - window // 2]
        x = X[k]
        if rfi_ratio > 0 and P[k] > np.float32(rfi_ratio) * m:
            x = x * math.sqrt(float(m) / float(P[k]))
            rfi += 1
        Y[k] = x * float(np.float32(math.sqrt(math.log(2) / float(m))))
    if len(zap):
        seconds = n * sample_us * 1e-6
        seed = int(np.array([dec[0]], dtype=np.float32).view(np.uint32)[0])
        sigma = math.sqrt(2.0) / 2 * math.sqrt(padding)
        for lower, upper in zap:
            for k in range(int(lower * seconds + 0.5), int(upper * seconds + 0.5) + 1):
                if window // 2 <= k < len(X) - window // 2:
                    r = math.sqrt(-2.0 * math.log(1.0 - _uniform(seed, 2 * k)))
                    a = 2 * math.pi * _uniform(seed, 2 * k + 1)
                    Y[k] = complex(float(np.float32(sigma * r * math.cos(a))), float(np.float32(sigma * r * math.sin(a))))
    y = np.fft.irfft(Y, n) * n / math.sqrt(n)
    return y[:samples], dec, rfi


@pytest.mark.parametrize("n,bits", [(4096, 4), (3 * 2048, 4), (4096, 8)])
def test_preprocess_matches_numpy_restatement(oracle, brp, n, bits):
    window = 64
    mean = 7.5 if bits == 4 else 0.0
    block, rfi_bins = brp.synthetic_dd(n, 64e-6, 3, mean, 2.0, bits, 200.0, 0.05, 0.1, (3600.0, 0.05, 1.0), n, 3, 50.0)
    out, dec, nrfi = oracle.preprocess(block, n, bits, 1.5, n, window, 30.0)
    ref, dec_np, rfi_np = _numpy_preprocess(block, n, bits, 1.5, n, window, 30.0)
    assert np.array_equal(dec, dec_np)
    assert nrfi == rfi_np == 3
    rms = float(np.sqrt(np.mean(ref ** 2)))
    assert np.max(np.abs(out - ref)) <= 1e-6 * rms
    # whitened noise: unit mean power per bin, so the series has unit variance (minus the
    # zeroed edge bins)
    assert 0.85 < float(np.var(out)) < 1.1


def test_uniform_matches_library_and_published_finaliser(oracle, brp):
    """The oracle's counter-based uniform (the zap noise source) equals the library's
    and a restatement from the published SplitMix64 finaliser, including counters whose
    products wrap."""
    for seed, counter in [(0, 0), (1, 1), (0x3F800000, 12345), (2 ** 32 - 1, 2 ** 40 + 7), (7, 2 ** 63 - 1)]:
        u = oracle.uniform(seed, counter)
        assert u == brp.load().brp_uniform(seed, counter) == _uniform(seed, counter)
        assert 0.0 <= u < 1.0


def test_preprocess_zap_list(oracle, brp):
    """The zap list (src/workload.cpp:76-86) against the numpy restatement: the listed
    ranges' bins carry the noise, the band edges stay zero, and the reference's errors
    for ranges past Nyquist and malformed ranges."""
    n, window, sample_us, padding = 4096, 64, 64.0, 1.0
    block, _ = brp.synthetic_dd(n, 64e-6, 5, 7.5, 2.0, 4, 200.0, 0.05, 0.1, (3600.0, 0.05, 1.0), n, 0, 50.0)
    seconds = n * sample_us * 1e-6
    # bins 40..50, one range inside the low edge (bins 0..4 < window / 2, zero) and one
    # single bin
    zap = [(40 / seconds, 50 / seconds), (0.0, 4.2 / seconds), (1000 / seconds, 1000.2 / seconds)]
    out, dec, _ = oracle.preprocess(block, n, 4, 1.5, n, window, 0.0, zap, sample_us, padding)
    ref, _, _ = _numpy_preprocess(block, n, 4, 1.5, n, window, 0.0, zap, sample_us, padding)
    plain, _, _ = oracle.preprocess(block, n, 4, 1.5, n, window, 0.0)
    rms = float(np.sqrt(np.mean(ref ** 2)))
    assert np.max(np.abs(out - ref)) <= 1e-6 * rms
    # Exactly the listed in-band bins changed: bins 40..50 and 1000 (the edge range 0..4
    # is zero either way).
    d = np.abs(np.fft.rfft(out.astype(np.float64) - plain.astype(np.float64)))
    changed = set(np.nonzero(d > 1e-3 * d.max())[0].tolist())
    assert changed == set(range(40, 51)) | {1000}
    with pytest.raises(RuntimeError, match="Nyquist"):
        oracle.preprocess(block, n, 4, 1.5, n, window, 0.0, [(0.0, (n / 2 + 1) / seconds)], sample_us, padding)
    for bad in ([(-1.0, 2.0)], [(3.0, 2.0)], [(math.nan, 2.0)]):
        with pytest.raises(RuntimeError, match="zap range"):
            oracle.preprocess(block, n, 4, 1.5, n, window, 0.0, bad, sample_us, padding)


def test_preprocess_rejects_bad_windows(oracle, brp):
    block, _ = brp.synthetic_dd(1024, 64e-6, 1, n=1024)
    for window in (0, 63, 1024):
        with pytest.raises(RuntimeError):
            oracle.preprocess(block, 1024, 4, 1.0, 1024, window, 0.0)

