"""GPU parity of both engines against the oracle, through the C ABI (libbrp_hip.so).

Every case prints its name before touching the GPU.  Sizes: the N = 65536 and
3 * 2^15 cases run the whole oracle; the N = 2^22 and 3 * 2^20 cases compare against
the oracle at full size (its FP64 FFT takes about a second there).
"""
import sys

import math

import numpy as np
import pytest

import parity

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def wl():
    import workloads

    return workloads


PLAIN_VARIANTS = {
    # the reference's selected configuration: packed, Stockham, radix-8 store, fused power, scatter
    "selected": dict(stockham=True, packed=True, radix8=True, radix8_store=True, fused_power=True, scatter=True),
    "radix4_gather": dict(stockham=True, packed=True, radix4=True),
    "full_complex_radix2": dict(stockham=True),
}


@pytest.mark.parametrize("variant", sorted(PLAIN_VARIANTS))
@pytest.mark.parametrize("scale", [1.0, 0.25])
def test_plain_small(brp, oracle, wl, variant, scale):
    w = wl.plain_small(threshold_scale=scale)
    p = w.templates[0]
    say(f"[gpu] plain N=65536 variant={variant} scale={scale}")
    with brp.GPU(p, w.x, 256, 0, 1, snapshot=True, **PLAIN_VARIANTS[variant]) as g:
        g.set_template(p)
        g.run()
        rep = parity.check_template(g.snapshot(0), oracle, w.x, p, f"plain/{variant}")
    say(rep)
    if scale < 1:
        assert rep["candidates"]["ref"] > 0


@pytest.mark.parametrize("scale", [1.0, 0.25])
@pytest.mark.parametrize("flags", [0, 6])
def test_fused_small(brp, oracle, wl, scale, flags):
    w = wl.plain_small(threshold_scale=scale, flags=flags)
    p = w.templates[0]
    say(f"[gpu] fused N=65536 scale={scale} flags={flags}")
    with brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, snapshot=True) as g:
        g.set_template(p)
        g.run()
        rep = parity.check_template(g.snapshot(0), oracle, w.x, p, "fused/small")
    say(rep)


@pytest.mark.parametrize("flags", [0, 6])
def test_fused_negative_phase(brp, oracle, wl, flags):
    """psi < 0: phases below zero take the general index form (wrap of a negative
    fraction), the other templates of the bank the non-negative form."""
    w = wl.plain_small(threshold_scale=0.3, flags=flags)
    p = w.templates[0].copy()
    p.psi = -1.3
    brp.resampling_steps(p)
    say(f"[gpu] fused N=65536 psi=-1.3 flags={flags}")
    with brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, snapshot=True) as g:
        g.set_template(p)
        g.run()
        snap = g.snapshot(0)
        rep = parity.check_template(snap, oracle, w.x, p, "fused/psi<0")
    say(rep)
    assert rep["candidates"]["ref"] > 0
    # The same template through the hot (no snapshot, graph) instance: same candidates.
    with brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, graph=True) as g:
        g.set_template(p)
        g.run()
        hot = g.candidates(0)
    assert np.array_equal(hot, snap.candidates)


@pytest.mark.parametrize("n", [1 << 13, 1 << 14, 1 << 15, 3 << 15, 1 << 17, 3 << 17, 3 << 18, 3 << 19, 1 << 20, 1 << 21,
                               3 << 21, 1 << 23, 3 << 22])
def test_fused_lengths(brp, oracle, wl, n):
    w = wl.plain_full(seed=5, n=n, threshold_scale=0.3)
    p = w.templates[0]
    say(f"[gpu] fused N={n}")
    with brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, snapshot=True) as g:
        g.set_template(p)
        g.run()
        rep = parity.check_template(g.snapshot(0), oracle, w.x, p, f"fused/N={n}")
    say(rep)


def test_plain_mixed_radix_small(brp, oracle, wl):
    w = wl.plain_full(seed=6, n=3 << 15, threshold_scale=0.3)
    p = w.templates[0]
    say("[gpu] plain N=3*2^15 selected")
    with brp.GPU(p, w.x, 256, 0, 1, snapshot=True, **PLAIN_VARIANTS["selected"]) as g:
        g.set_template(p)
        g.run()
        rep = parity.check_template(g.snapshot(0), oracle, w.x, p, "plain/3*2^15")
    say(rep)


@pytest.mark.parametrize("variant", ["selected", "full_complex_radix2"])
@pytest.mark.parametrize("n", [64, 96, 128, 192, 1 << 10, 3 << 10, 1 << 12])
def test_plain_smallest_lengths(brp, oracle, wl, n, variant):
    """The reference accepts n in {2^k, 3 * 2^k}, 64 <= n (src/science.cpp:46-65); below
    the fused engine's shapes the plain engine serves those lengths at the same bar."""
    w = wl.plain_full(seed=11, n=n, threshold_scale=0.3)
    p = w.templates[0]
    say(f"[gpu] plain N={n} variant={variant}")
    with brp.GPU(p, w.x, 256, 0, 1, snapshot=True, **PLAIN_VARIANTS[variant]) as g:
        g.set_template(p)
        g.run()
        rep = parity.check_template(g.snapshot(0), oracle, w.x, p, f"plain/N={n}/{variant}")
    say(rep)


@pytest.mark.parametrize("n", [64, 3 << 12, 1 << 12])
def test_fused_rejects_lengths_below_its_shapes(brp, wl, n):
    w = wl.plain_full(seed=11, n=n, threshold_scale=0.3)
    say(f"[gpu] fused N={n} (rejected)")
    with pytest.raises(RuntimeError, match="unsupported series length"):
        brp.GPU(w.templates[0], w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED)


@pytest.mark.parametrize("engine", ["fused", "plain"])
@pytest.mark.parametrize("n,unpadded,tau", [(1 << 16, 49152, 0.05), (3 << 15, 60000, 0.05), (1 << 16, 65536, 0.0),
                                            (1 << 17, 100000, 0.0)])
def test_padding_and_zero_delay(brp, oracle, wl, engine, n, unpadded, tau):
    """An observation shorter than the transform (a long mean pad, which the fused engine
    adds by linearity in pass 2) and an orbit of zero delay (every index j = i)."""
    w = wl.plain_full(seed=13, n=n, threshold_scale=0.3)
    x = np.ascontiguousarray(w.x[:unpadded])
    p = brp.parameters(n, unpadded, tau, wl.INJECTED[0], wl.INJECTED[2], wl.DT)
    for h in range(5):
        p.threshold[h] *= 0.3
    say(f"[gpu] {engine} N={n} unpadded={unpadded} tau={tau} steps={p.steps}")
    kw = dict(packed=True, engine=brp.ENGINE_FUSED) if engine == "fused" else PLAIN_VARIANTS["selected"]
    with brp.GPU(p, x, 256, 0, 1, snapshot=True, **kw) as g:
        g.set_template(p)
        g.run()
        snap = g.snapshot(0)
        rep = parity.check_template(snap, oracle, x, p, f"{engine}/N={n}/U={unpadded}/tau={tau}")
    if tau == 0.0:
        assert np.array_equal(snap.index[:p.steps], np.arange(p.steps))
    say(rep)


@pytest.mark.parametrize("engine", ["fused", "plain"])
def test_full_size_2p22(brp, oracle, wl, engine):
    w = wl.plain_full(seed=42)
    p = w.templates[0]
    say(f"[gpu] {engine} N=2^22 spin={w.spin_hz:.3f} Hz")
    kw = dict(packed=True, engine=brp.ENGINE_FUSED) if engine == "fused" else PLAIN_VARIANTS["selected"]
    with brp.GPU(p, w.x, 256, 0, 1, snapshot=True, **kw) as g:
        g.set_template(p)
        g.run()
        snap = g.snapshot(0)
        rep = parity.check_template(snap, oracle, w.x, p, f"{engine}/2^22")
        # The injected pulsar is found by the matching template.
        assert len(snap.candidates) > 0
        g.run()
        again = g.snapshot(0)
        assert np.array_equal(snap.power, again.power) and np.array_equal(snap.summed, again.summed)
    say(rep)


# (series length N, pass-1 length C, rows per pass-1 block S1, column pairs per pass-2
# block S2) of every non-default shape; N/2 = C x R.
ALTERNATIVE_SPLITS = [
    (1 << 20, 1024, 16, 16), (1 << 20, 4096, 4, 16), (1 << 20, 2048, 4, 16),
    (1 << 21, 2048, 8, 16), (1 << 21, 1024, 16, 8), (1 << 21, 4096, 2, 16),
    (1 << 22, 2048, 8, 8), (1 << 22, 1024, 16, 4), (1 << 22, 4096, 2, 8), (1 << 22, 4096, 2, 16),
    (1 << 22, 4096, 4, 8), (1 << 22, 2048, 4, 4),
    (1 << 23, 2048, 8, 4), (1 << 23, 4096, 2, 4),
    (3 << 19, 1536, 8, 16), (3 << 19, 768, 16, 8), (3 << 20, 1536, 8, 8), (3 << 20, 768, 16, 4),
    (3 << 21, 1536, 8, 4),
    # 3 * 2^k with the factor 3 in pass 1 (C = 3 * 2^j, 24 elements per thread) or with the
    # pass-2 length 3 * 2^j at 16 column pairs per block
    (3 << 15, 384, 32, 16), (3 << 16, 768, 16, 16), (3 << 17, 768, 16, 16), (3 << 18, 1536, 8, 16),
    (3 << 19, 3072, 2, 16), (3 << 20, 3072, 2, 16), (3 << 21, 3072, 2, 8), (3 << 22, 3072, 2, 4),
    (3 << 19, 4096, 2, 16), (3 << 20, 4096, 2, 16), (3 << 21, 4096, 2, 8),
    (1 << 22, 4096, 1, 4), (1 << 22, 4096, 2, 4), (1 << 22, 4096, 1, 8), (1 << 22, 2048, 2, 2),
    (1 << 21, 4096, 1, 8), (1 << 22, 4096, 4, 16),
    # eight elements per thread (radix-8 plans): (N, C, S1, S2, T1, T2)
    (1 << 22, 4096, 2, 4, 1024, 256), (1 << 22, 4096, 1, 4, 512, 256), (1 << 22, 4096, 2, 4, 512, 512),
    (1 << 22, 4096, 1, 4, 512, 512),
    # 4096-point rows in 512-thread blocks, 256-thread pass-2 blocks
    (1 << 22, 4096, 2, 4, 512, 256),
    # 2048-point rows, eight elements per thread (radix 8, 8, 8, 4)
    (1 << 22, 2048, 2, 4, 512, 512),
    # the earlier defaults of 2^20, 2^21, 3 * 2^19 and 3 * 2^20
    (1 << 20, 2048, 8, 16, 1024, 512), (1 << 21, 4096, 4, 16, 1024, 512), (3 << 19, 2048, 4, 8, 512, 256),
    (3 << 20, 4096, 2, 8, 512, 256)]


@pytest.mark.parametrize("shape", ALTERNATIVE_SPLITS)
def test_fused_alternative_splits(brp, oracle, wl, shape):
    """Every non-default four-step shape the autotuner may pick meets the same bar."""
    n, c, s1, s2 = shape[:4]
    t1, t2 = shape[4:] if len(shape) > 4 else (0, 0)
    w = wl.plain_full(seed=7, n=n, threshold_scale=0.3)
    p = w.templates[0]
    say(f"[gpu] fused N={n} split C={c} S1={s1} S2={s2} T1={t1} T2={t2}")
    with brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, snapshot=True, split_c=c,
                 split_s1=s1, split_s2=s2, split_t1=t1, split_t2=t2) as g:
        assert f"C={c} " in g.device_info() and f"S1={s1} S2={s2}" in g.device_info()
        if t1:
            assert f"T1={t1} T2={t2}" in g.device_info()
        g.set_template(p)
        g.run()
        rep = parity.check_template(g.snapshot(0), oracle, w.x, p, f"fused/N={n}/C={c}/S1={s1}/S2={s2}/T={t1},{t2}")
    say(rep)


def test_autotune_picks_a_correct_configuration(brp, wl):
    """The tuner reports every (split, templates-per-launch) candidate and its pick gives
    the same candidates as the default configuration."""
    w = wl.bank(count=8, n=1 << 20)
    say("[gpu] autotune N=2^20, 8 templates")
    best, report = brp.autotune(w.templates, w.x, reps=2)
    say(best, report)
    shapes = {(r["C"], r["S1"], r["S2"], r["T1"], r["T2"]) for r in report}
    # every shape x (group, lanes) in (2, 1), (2, 2), (4, 1), (4, 2), (8, 1)
    assert len(report) == 5 * len(shapes) and len(shapes) >= 3
    assert {(r["group"], r["lanes"]) for r in report} == {(2, 1), (2, 2), (4, 1), (4, 2), (8, 1)}
    assert {s[0] for s in shapes} == {4096, 2048, 1024}
    assert (best["split_c"], best["split_s1"], best["split_s2"], best["split_t1"], best["split_t2"]) in shapes
    assert best["group"] in {2, 4, 8}
    assert best["lanes"] in {1, 2}
    assert all(r["us_per_template"] > 0 for r in report)
    out = []
    for kw in ({}, best):
        with brp.GPU(w.templates[0], w.x, 256, 0, 8, packed=True, engine=brp.ENGINE_FUSED, graph=True, **kw) as g:
            g.set_templates(w.templates)
            g.run()
            out.append([g.candidates(s) for s in range(8)])
    for p, a, b in zip(w.templates, *out):
        cc = parity.compare_candidates(b, a, p.threshold)
        assert not cc["mismatches"], cc


def test_autotune_with_precision_variant(brp, wl):
    """With the split-FP16 variant the tuner times only the shapes that have the radix-16
    stages it asks for, and its pick runs with the variant."""
    w = wl.bank(count=4, n=1 << 20)
    say("[gpu] autotune N=2^20, 4 templates, precision=1 and precision=2")
    for prec in (1, 2):
        best, report = brp.autotune(w.templates, w.x, {"precision": prec, "lanes": 1}, reps=1)
        say(prec, best, report)
        assert report and all(r["us_per_template"] > 0 for r in report)
        if prec & 2:  # pass-1 shapes compiled with the variant: 2 rows of 2048 in 256 threads
            assert {(r["C"], r["S1"], r["T1"]) for r in report} == {(2048, 2, 256)}
        with brp.GPU(w.templates[0], w.x, 256, 0, 4, packed=True, engine=brp.ENGINE_FUSED, precision=prec,
                     **best) as g:
            g.set_templates(w.templates)
            g.run()


def test_full_size_mixed_radix(brp, oracle, wl):
    w = wl.mixed_radix(count=3)
    say("[gpu] fused N=3*2^20, 3 templates in one launch")
    with brp.GPU(w.templates[0], w.x, 256, 0, 3, packed=True, engine=brp.ENGINE_FUSED, snapshot=True) as g:
        g.set_templates(w.templates)
        g.run()
        for s, p in enumerate(w.templates):
            rep = parity.check_template(g.snapshot(s), oracle, w.x, p, f"fused/3*2^20/slot{s}")
            say(rep)


def test_fused_batch_slots_match_single_runs(brp, wl):
    w = wl.bank(count=6, n=1 << 20)
    say("[gpu] fused batch of 6 (group 4, graph) vs single-template runs, N=2^20")
    with brp.GPU(w.templates[0], w.x, 256, 0, 6, packed=True, engine=brp.ENGINE_FUSED, graph=True, group=4) as g:
        g.set_templates(w.templates)
        g.run()
        batched = [g.candidates(s) for s in range(6)]
        g.run()
        assert all(np.array_equal(a, g.candidates(s)) for s, a in enumerate(batched))
    with brp.GPU(w.templates[0], w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED) as g:
        for s, p in enumerate(w.templates):
            g.set_template(p)
            g.run()
            assert np.array_equal(g.candidates(0), batched[s])


@pytest.mark.parametrize("flags", [0, 6])
def test_fused_random_orbits_index_exact(brp, oracle, wl, flags):
    """48 seeded orbits in one batch (P log-uniform in [660, 10000] s, tau in [0, 0.3] s):
    the first 24 with psi in [0, 2 pi) (launch groups that take the non-negative index
    form), the rest with psi in (-2 pi, 2 pi) (the general form).  Every slot's resample
    indices equal the oracle's bit for bit, and its power, sums and candidates meet the bar."""
    n, count = 1 << 16, 48
    rng = np.random.default_rng(2026 + flags)
    w = wl.plain_small(threshold_scale=0.5)
    templates = []
    for k in range(count):
        period = float(np.exp(rng.uniform(np.log(660.0), np.log(10000.0))))
        tau = float(rng.uniform(0.0, 0.3))
        psi = float(rng.uniform(0.0, 2 * np.pi) if k < count // 2 else rng.uniform(-2 * np.pi, 2 * np.pi))
        p = brp.parameters(n, n, tau, period, psi, wl.DT)
        if flags:
            p.flags = flags
            p.s0 = float(np.float32(np.float32(p.tau) * np.float32(math.sin(np.float32(p.psi)))) * np.float32(p.step_inv))
            brp.resampling_steps(p)
        for h in range(5):
            p.threshold[h] *= 0.5
        templates.append(p)
    say(f"[gpu] fused N=2^16, {count} random orbits in one batch, flags={flags}")
    with brp.GPU(templates[0], w.x, 256, 0, count, packed=True, engine=brp.ENGINE_FUSED, snapshot=True,
                 group=8) as g:
        g.set_templates(templates)
        g.run()
        cands = 0
        for s, p in enumerate(templates):
            rep = parity.check_template(g.snapshot(s), oracle, w.x, p, f"fused/random orbit {s}/flags={flags}")
            cands += rep["candidates"]["gpu"]
    say({"templates": count, "candidates": cands})


@pytest.mark.parametrize("graph", [True, False])
def test_fused_two_lanes_match_one(brp, wl, graph):
    """Two lanes (odd groups on a second stream with their own buffers) and untimed
    graphs give the same candidates as one lane, on repeated runs."""
    w = wl.bank(count=14, n=1 << 20)
    say(f"[gpu] fused batch of 14, group 2/3, lanes 1 vs 2, graph={graph}, N=2^20")
    ref = None
    for lanes, group, untimed in ((1, 2, False), (2, 2, False), (2, 3, True), (2, 2, True)):
        with brp.GPU(w.templates[0], w.x, 256, 0, 14, packed=True, engine=brp.ENGINE_FUSED, graph=graph,
                     group=group, lanes=lanes, untimed=untimed) as g:
            g.set_templates(w.templates)
            for _ in range(3):
                g.run()
                got = [g.candidates(s) for s in range(14)]
                if ref is None:
                    ref = got
                assert all(np.array_equal(a, b) for a, b in zip(ref, got)), (lanes, group, untimed)
    assert sum(len(c) for c in ref) > 0


def test_sampled_kernel_timing(brp, wl):
    """timing_stride n: events on every n-th launch group; kernel_ms reports those launches
    and the templates they covered, and the candidates do not change."""
    w = wl.bank(count=14, n=1 << 20)
    say("[gpu] timing stride 1 vs 3, batch 14, group 2, graph")
    got = []
    for stride in (1, 3):
        with brp.GPU(w.templates[0], w.x, 256, 0, 14, packed=True, engine=brp.ENGINE_FUSED, graph=True, group=2,
                     timing_stride=stride) as g:
            g.set_templates(w.templates)
            g.kernel_ms()
            for _ in range(2):
                g.run()
            km, launches, templates = g.kernel_ms()
            got.append([g.candidates(s) for s in range(14)])
        assert launches == 2 * len(range(0, 7, stride)) and templates == 2 * 2 * len(range(0, 7, stride))
        assert all(v > 0 for v in (km[0], km[1], km[3]))
    assert all(np.array_equal(a, b) for a, b in zip(*got))


def _oracle_bank(oracle, x, templates, orbits, rows=100):
    top = oracle.OToplist(rows)
    for p, o in zip(templates, orbits):
        q = oracle.as_oparams(p)
        top.thresholds(q)
        res, idx = oracle.resample(x, q)
        p32, _ = oracle.power(res)
        cands = oracle.select(oracle.harmonic(p32, q), q)
        top.retain(cands, o, q)
    return top


def test_bank_search_device_lengths_equal_host_parameters(brp, wl):
    """brp_hip_bank_search on the fused engine leaves the resampled lengths to the device
    (resampling_steps in the launch stream of each batch).  The parameter blocks it ran equal
    brp_search_parameters byte for byte, apart from the thresholds the toplist raised: 1024 seeded bank orbits at N = 2^20 and 2^22 plus
    edge orbits (no delay, psi < 0 and > 2 pi, a tiny delay, a delay of 32 s, and delays of
    40, 60 and 150 s, which exceed the jump start's error bound and keep the host's scan; all
    with every resampled index inside the series)."""
    edge = [(3600.0, 0.0, 1.0), (660.0, 0.3, 6.2831), (5000.0, 0.2, -1.0), (5000.0, 0.2, 7.5), (700.0, 1e-6, 3.0),
            (20000.0, 32.0, 2.0), (40000.0, 40.0, 0.5), (80000.0, 60.0, 0.3), (200000.0, 150.0, 0.1)]
    for n in (1 << 20, 1 << 22):
        w = wl.bank(count=1024 - len(edge), n=n)
        orbits = edge + list(w.orbits)
        say(f"[gpu] bank_search lengths on the device, N={n}, {len(orbits)} orbits")
        with brp.GPU(w.templates[0], w.x, 256, 0, 128, packed=True, engine=brp.ENGINE_FUSED, graph=True,
                     untimed=True) as g:
            ran, counts, _ = g.bank_search(orbits, n, wl.SAMPLE_US, brp.Toplist(100))
        for o, p in zip(orbits, ran):
            want = brp.search_parameters(n, n, wl.SAMPLE_US, o)
            # The blocks come back as run: thresholds raised by the toplist, never lowered.
            assert all(p.threshold[h] >= want.threshold[h] for h in range(5)), o
            for h in range(5):
                want.threshold[h] = p.threshold[h]
            assert bytes(p) == bytes(want), o
        assert counts.sum() > 0


def test_bank_toplist_identical(brp, oracle, wl):
    count, n = 24, 1 << 20
    w = wl.bank(count=count, n=n)
    say(f"[gpu] bank {count} x N={n}: fused batched (batch 8) vs plain batch 1 vs oracle loop")
    tops = {}
    for name, kw in {"fused": dict(packed=True, engine=brp.ENGINE_FUSED, graph=True, batch=8),
                     "plain": dict(engine=brp.ENGINE_PLAIN, batch=1, **PLAIN_VARIANTS["selected"])}.items():
        batch = kw.pop("batch")
        top = brp.Toplist(100)
        with brp.GPU(w.templates[0], w.x, 256, 0, batch, **kw) as g:
            g.bank_search(w.orbits, n, 64.0, top)
        tops[name] = top
    ref = _oracle_bank(oracle, w.x, w.templates, w.orbits)
    refrows = [(r.bin, r.harmonics, r.period) for r in ref.state]
    for name, top in tops.items():
        rows = [(r.bin, r.harmonics, r.period) for r in top.table()]
        assert rows == refrows, name
        pw = np.array([r.power for r in top.table()])
        rp = np.array([r.power for r in ref.state])
        assert np.allclose(pw, rp, rtol=1e-5, atol=0), name
        assert [r.bin for r in top.final()] == [r.bin for r in ref.final()], name
    assert sum(r.harmonics > 0 for r in ref.state) > 0


def test_fused_uneven_batch_parity(brp, oracle, wl):
    """Templates of three series lengths (2^k and 3 * 2^k) in one batch: every slot meets
    the full parity bar against the oracle on its own series."""
    u = wl.uneven(count=9, lengths=(1 << 17, 3 << 16, 1 << 18))
    say(f"[gpu] fused uneven batch of 9 over N={[s[0] for s in u.series]}, snapshots")
    xs = {n: x for n, x, _ in u.series}
    first = u.templates[0]
    with brp.GPU(first, xs[first.n], 256, 0, 9, packed=True, engine=brp.ENGINE_FUSED, snapshot=True) as g:
        for n, x, _ in u.series:
            if n != first.n:
                g.add_series(brp.search_parameters(n, n, wl.SAMPLE_US, wl.INJECTED), x)
        g.set_templates(u.templates)
        g.run()
        for s, p in enumerate(u.templates):
            rep = parity.check_template(g.snapshot(s), oracle, xs[p.n], p, f"uneven/slot{s}/N={p.n}")
            say(rep)


@pytest.mark.parametrize("lanes", [1, 2])
def test_bank_uneven_toplists(brp, oracle, wl, lanes):
    """A bank split over three lengths in one launch schedule (mixed batches, graph): each
    series' toplist equals the oracle's one-template-at-a-time loop over its templates."""
    u = wl.uneven(count=40, lengths=(1 << 18, 3 << 17, 1 << 19))
    say(f"[gpu] bank of 40 over N={[s[0] for s in u.series]}, batch 16, group 4, lanes {lanes}")
    first = u.templates[0]
    xs = {n: x for n, x, _ in u.series}
    order = [n for n, _, _ in u.series]
    reg = [first.n] + [n for n in order if n != first.n]  # registration order = toplist order
    tops = [brp.Toplist(100) for _ in reg]
    with brp.GPU(first, xs[first.n], 256, 0, 16, packed=True, engine=brp.ENGINE_FUSED, graph=True, group=4,
                 lanes=lanes) as g:
        for n in reg[1:]:
            g.add_series(brp.search_parameters(n, n, wl.SAMPLE_US, wl.INJECTED), xs[n])
        counts, _ = g.bank_search_templates(u.templates, u.orbits, tops)
    assert counts.sum() > 0
    for k, n in enumerate(reg):
        sel = [i for i, p in enumerate(u.templates) if p.n == n]
        ref = _oracle_bank(oracle, xs[n], [u.templates[i] for i in sel], [u.orbits[i] for i in sel])
        rows = [(r.bin, r.harmonics, r.period) for r in tops[k].table()]
        assert rows == [(r.bin, r.harmonics, r.period) for r in ref.state], n
        pw = np.array([r.power for r in tops[k].table()])
        assert np.allclose(pw, np.array([r.power for r in ref.state]), rtol=1e-5, atol=0), n
        assert [r.bin for r in tops[k].final()] == [r.bin for r in ref.final()], n


@pytest.mark.parametrize("ties", [False, True])
def test_device_retain_equals_host_on_random_streams(brp, ties):
    """Toplist::retain on the device (brp_hip_toplist_retain_device) against the host
    retain on random candidate streams: repeated bins, tie-heavy powers, rows per level
    from 3 to 1000 (the stable single move of equal powers included).  Within a template
    the (level, bin) pairs are distinct, as the compaction produces them."""
    rng = np.random.default_rng(40 + ties)
    n = 1 << 20
    for rows in (3, 7, 100, 1000):
        host, dev = brp.Toplist(rows), brp.Toplist(rows)
        templates, orbits, lists = [], [], []
        for t in range(80):
            orbit = (float(rng.uniform(660, 10000)), float(rng.uniform(0, 0.3)), float(rng.uniform(0, 6.28)))
            p = brp.search_parameters(n, n, 64.0, orbit)
            k = int(rng.integers(0, 40))
            levels = rng.integers(0, 5, k)
            bins = rng.integers(p.window, p.fundamental, k)
            bins[: k // 3] = rng.integers(p.window, p.window + 5, k // 3)
            if ties:
                powers = np.float32([100.0 * int(rng.integers(1, 4)) for _ in levels])
            else:
                powers = np.float32([p.threshold[l] * rng.uniform(0.9, 3.0) for l in levels])
            seen, cands = set(), []
            for b, lv, pw in zip(bins, levels, powers):
                if (int(lv), int(b)) not in seen:
                    seen.add((int(lv), int(b)))
                    cands.append(brp.Candidate(int(b), int(lv), float(pw), 0))
            templates.append(p)
            orbits.append(orbit)
            lists.append(list(reversed(cands)))  # any order within a template
            q = p.copy()
            host.thresholds(q)
            host.retain(sorted(cands, key=lambda c: (c.level, c.bin)), orbit, q)
        say(f"[gpu] device retain rows={rows} ties={ties} templates=80")
        brp.toplist_retain_device(dev, templates, orbits, lists)
        mine = [(r.bin, r.power, r.period, r.tau, r.psi, r.harmonics) for r in dev.table()]
        ref = [(r.bin, r.power, r.period, r.tau, r.psi, r.harmonics) for r in host.table()]
        assert mine == ref, rows


@pytest.mark.parametrize("rows,lanes", [(100, 1), (1000, 1), (100, 2)])
def test_bank_search_device_toplist_equals_host(brp, wl, rows, lanes):
    """The bank with its toplists on the device (thresholds raised in the launch stream,
    retain overlapping the next batch on its own stream) against the host bank loop:
    identical toplists for every series of an uneven bank, and identical candidate counts
    (both freeze a batch's thresholds at the toplist two batches back)."""
    u = wl.uneven(count=48, lengths=(1 << 18, 3 << 17, 1 << 19))
    first = u.templates[0]
    xs = {n: x for n, x, _ in u.series}
    reg = [first.n] + [n for n, _, _ in u.series if n != first.n]
    say(f"[gpu] bank_search_device 48 templates over {reg}, rows={rows}, lanes={lanes}")
    out = {}
    for mode in ("host", "device"):
        tops = [brp.Toplist(rows) for _ in reg]
        with brp.GPU(first, xs[first.n], 256, 0, 16, packed=True, engine=brp.ENGINE_FUSED, graph=True, group=4,
                     lanes=lanes) as g:
            for n in reg[1:]:
                g.add_series(brp.search_parameters(n, n, wl.SAMPLE_US, wl.INJECTED), xs[n])
            if mode == "host":
                counts, _ = g.bank_search_templates(u.templates, u.orbits, tops)
            else:
                counts, _ = g.bank_search_device(u.templates, u.orbits, tops)
        out[mode] = ([[(r.bin, r.power, r.period, r.tau, r.psi, r.harmonics) for r in t.table()] for t in tops],
                     counts)
    assert out["host"][0] == out["device"][0]
    assert out["device"][1].sum() > 0
    assert np.array_equal(out["host"][1], out["device"][1])


@pytest.mark.parametrize("n,bits", [(1 << 16, 4), (3 << 15, 4), (1 << 16, 8)])
def test_preprocess_parity(brp, oracle, n, bits):
    """Device decode + whitening + RFI replacement against the FP64 oracle: decoded
    samples bit for bit, the same RFI bins replaced, rel-L2 <= 1e-5 on the series."""
    mean = 7.5 if bits == 4 else 0.0
    block, injected = brp.synthetic_dd(n, 64e-6, 11, mean, 2.0, bits, 200.0, 0.05, 0.1, (3600.0, 0.05, 1.0), n, 10,
                                       50.0)
    say(f"[gpu] preprocess N={n} bits={bits}")
    out, dec, rfi = brp.preprocess(block, n, bits, 1.5, n, 1000, 30.0)
    ref, rdec, rrfi = oracle.preprocess(block, n, bits, 1.5, n, 1000, 30.0)
    assert np.array_equal(dec, rdec)
    assert rfi == rrfi == 10
    rel = float(np.linalg.norm(out.astype(np.float64) - ref) / np.linalg.norm(ref))
    say({"rel_l2": rel, "max_abs": float(np.max(np.abs(out - ref)))})
    assert rel <= 1e-5


def test_preprocess_zap_parity(brp, oracle):
    """The zap list on the device against the oracle: the same decoded samples and a
    series within rel-L2 1e-5, with ranges inside the band, across its low edge and a
    single bin, and the reference's error for a range past Nyquist."""
    n, sample_us, padding = 1 << 16, 64.0, 1.25
    block, _ = brp.synthetic_dd(n, 64e-6, 13, 7.5, 2.0, 4, 200.0, 0.05, 0.1, (3600.0, 0.05, 1.0), n, 10, 50.0)
    seconds = n * sample_us * 1e-6
    zap = [(37.0, 41.5), (0.0, 2.0), (1000.0, 1000.0), (1500.0, 1530.25)]
    say(f"[gpu] preprocess zap N={n} ranges={zap}")
    out, dec, rfi = brp.preprocess(block, n, 4, 1.5, n, 1000, 30.0, zap, sample_us, padding)
    ref, rdec, rrfi = oracle.preprocess(block, n, 4, 1.5, n, 1000, 30.0, zap, sample_us, padding)
    assert np.array_equal(dec, rdec) and rfi == rrfi
    rel = float(np.linalg.norm(out.astype(np.float64) - ref) / np.linalg.norm(ref))
    say({"rel_l2": rel, "seconds": seconds})
    assert rel <= 1e-5
    with pytest.raises(RuntimeError, match="Nyquist"):
        brp.preprocess(block, n, 4, 1.5, n, 1000, 30.0, [(0.0, (n / 2 + 1) / seconds)], sample_us, padding)


def test_uneven_workunit_pipeline(brp, oracle, wl):
    """The uneven workunit end to end at reduced sizes: 4-bit DD blocks of three lengths
    with pulsars and RFI, preprocessed on the device, then one mixed bank in one launch
    schedule; each series' toplist equals the oracle loop on the same preprocessed series."""
    lengths = (1 << 17, 3 << 16, 1 << 18)
    xs = []
    for k, n in enumerate(lengths):
        spin = wl.spin_for(42 + 10 * k)
        block, _ = brp.synthetic_dd(n, wl.DT, 42 + 10 * k, 7.5, 2.0, 4, spin, 0.05, 0.1, wl.INJECTED, n, 10, 50.0)
        x, _, rfi = brp.preprocess(block, n, 4, 1.0, n, 1000, 30.0)
        assert rfi == 10
        xs.append(x)
    u = wl.uneven(count=36, lengths=lengths, xs=xs)
    say(f"[gpu] uneven workunit: 36 templates over N={list(lengths)} (4-bit, whitened, RFI replaced)")
    first = u.templates[0]
    byn = {n: x for n, x, _ in u.series}
    reg = [first.n] + [n for n in lengths if n != first.n]
    tops = [brp.Toplist(100) for _ in reg]
    with brp.GPU(first, byn[first.n], 256, 0, 16, packed=True, engine=brp.ENGINE_FUSED, graph=True, group=4) as g:
        for n in reg[1:]:
            g.add_series(brp.search_parameters(n, n, wl.SAMPLE_US, wl.INJECTED), byn[n])
        counts, _ = g.bank_search_templates(u.templates, u.orbits, tops)
    assert counts.sum() > 0
    for k, n in enumerate(reg):
        sel = [i for i, p in enumerate(u.templates) if p.n == n]
        ref = _oracle_bank(oracle, byn[n], [u.templates[i] for i in sel], [u.orbits[i] for i in sel])
        assert [(r.bin, r.harmonics, r.period) for r in tops[k].table()] == \
            [(r.bin, r.harmonics, r.period) for r in ref.state], n
        assert [r.bin for r in tops[k].final()] == [r.bin for r in ref.final()], n


@pytest.mark.parametrize("n", [3 << 20, 1 << 22])
def test_split_fp16_mfma_variant(brp, oracle, wl, n):
    """precision = 1 / 2 / 3: the radix-16 stages of pass 2, of pass 1 (the transform's first
    stages) or of both as split-FP16 MFMA GEMMs.  Indices and the resampled series are
    unchanged; every placement keeps the power within the parity bar (1e-5 relative of the
    FP64 oracle), and pass 2's within the variant's 2e-6 target.  Pass 1's inputs are the raw
    resampled samples, whose 22-bit split is scaled per wave by the largest of them, so its
    placement does not hold 2e-6 at every bin (up to about 6e-6 near DC at 3 * 2^20); the
    max / RMS errors are reported beside the FP32 path's.  At 3 * 2^20 no compiled shape has
    radix-16 stages in both passes."""
    w = wl.mixed_radix(count=2, n=n) if n % 3 == 0 else wl.plain_full(seed=9, n=n, threshold_scale=0.3)
    p = w.templates[0]
    errs = {}
    precisions = (0, 1, 2, 3) if n == 1 << 22 else (0, 1, 2)
    for prec in precisions:
        say(f"[gpu] split-FP16 MFMA variant N={n} precision={prec}")
        with brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, snapshot=True, precision=prec) as g:
            info = g.device_info()
            g.set_template(p)
            g.run()
            rep = parity.check_template(g.snapshot(0), oracle, w.x, p, f"fused/N={n}/precision={prec}")
        errs[prec] = rep["power"]
        if prec & 1:
            assert " R=1024 " in info or " R=2048 " in info  # a shape with radix-16 pass-2 stages
        if prec & 2:
            assert " C=2048 " in info or " C=4096 " in info  # a pass-1 shape compiled with them
    say({prec: errs[prec] for prec in precisions})
    assert errs[1]["max_rel"] <= 2e-6  # the variant's target against the FP64 oracle
    for prec in precisions[1:]:
        assert errs[prec]["max_rel"] <= 1e-5, prec  # the parity bar
    if n % 3 == 0:
        with pytest.raises(RuntimeError, match="radix-16"):
            brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, precision=3)


def test_split_fp16_rejected_without_radix16_stages(brp, wl):
    """N = 2^21 with C = 2048 leaves R = 512 (radix 8 only): the variant has no stage to run."""
    w = wl.plain_full(seed=3, n=1 << 21)
    p = w.templates[0]
    with pytest.raises(RuntimeError, match="radix-16"):
        brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, precision=1, split_c=2048)
    with pytest.raises(RuntimeError, match="fused engine"):
        brp.GPU(p, w.x, 256, 0, 1, packed=True, precision=1)


def test_fp64_phase_variant(brp, oracle, wl):
    """The FP64-phase variant rounds the exact orbit model; its index differences from the
    FP32 LUT rule (the parity mode) are counted and reported, and its spectrum is checked
    against the FP64 FFT of its own resampled series."""
    w = wl.bank(count=2, n=1 << 20)
    p = w.templates[0]
    say("[gpu] fused FP64-phase variant, N=2^20")
    with brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, snapshot=True, fp64_phase=True) as g:
        g.set_template(p)
        g.run()
        snap = g.snapshot(0)
    steps = p.steps
    i = np.arange(steps, dtype=np.float64)
    tau_dt = float(np.float32(p.tau)) / float(np.float32(p.dt))
    s0_dt = float(np.float32(p.tau)) * np.sin(float(np.float32(p.psi))) / float(np.float32(p.dt))
    off = tau_dt * np.sin(float(np.float32(p.omega)) * (i * float(np.float32(p.dt))) + float(np.float32(p.psi))) - s0_dt
    resid = i - off - snap.index
    assert np.max(np.abs(resid)) <= 0.5 + 1e-6
    _, idx32 = oracle.resample(w.x, p)
    differ = int(np.count_nonzero(idx32 != snap.index))
    say({"fp64_phase_index_differences": differ, "of": steps})
    assert differ < 0.05 * steps
    assert np.array_equal(snap.resampled[:steps], w.x[snap.index])
    _, p64 = oracle.power(snap.resampled)
    assert parity.power_error(snap.power, p64)["ok"]


def test_candidate_overflow_raises(brp, wl):
    w = wl.plain_small(threshold_scale=0.2)
    p = w.templates[0]
    p.capacity = 4
    say("[gpu] overflow: capacity 4 with low thresholds")
    for kw in (dict(packed=True, engine=brp.ENGINE_FUSED), dict(PLAIN_VARIANTS["selected"])):
        with brp.GPU(p, w.x, 256, 0, 1, **kw) as g:
            g.set_template(p)
            g.run()
            with pytest.raises(RuntimeError, match="overflow"):
                g.candidates(0)


def test_device_retain_reports_invalid_candidates(brp):
    """The device retain checks every candidate as the host does: a level past 4, a bin
    outside [W, F) or a non-finite power is an error, and so is a level with more
    candidates in one template than the retain can sort."""
    n = 1 << 20
    orbit = (3600.0, 0.05, 1.0)
    p = brp.search_parameters(n, n, 64.0, orbit)
    bad = {
        "level 5": [brp.Candidate(p.window + 10, 5, 100.0, 0)],
        "bin below W": [brp.Candidate(p.window - 1, 0, 100.0, 0)],
        "bin at F": [brp.Candidate(p.fundamental, 1, 100.0, 0)],
        "NaN power": [brp.Candidate(p.window + 3, 2, float("nan"), 0)],
    }
    for name, cands in bad.items():
        say(f"[gpu] device retain rejects {name}")
        with pytest.raises(RuntimeError, match="invalid compact candidate"):
            brp.toplist_retain_device(brp.Toplist(100), [p], [orbit], [cands])
    many = [brp.Candidate(p.window + i, 0, 100.0 + i, 0) for i in range(5000)]
    say("[gpu] device retain rejects 5000 candidates of one level")
    with pytest.raises(RuntimeError, match="more than 4096"):
        brp.toplist_retain_device(brp.Toplist(100), [p], [orbit], [many])


def test_bank_search_device_reports_overflow(brp, wl):
    """A template whose candidates overflow its capacity makes the device-toplist bank
    search fail, as the host loop does, instead of retaining a partial list."""
    w = wl.bank(count=16, n=1 << 16)
    templates = [t.copy() for t in w.templates]
    for t in templates:
        t.capacity = 4
        for h in range(5):
            t.threshold[h] *= 0.2
    say("[gpu] device-toplist bank search with capacity 4 and low thresholds")
    with brp.GPU(templates[0], w.x, 256, 0, 16, packed=True, engine=brp.ENGINE_FUSED, graph=True) as g:
        with pytest.raises(RuntimeError, match="overflow"):
            g.bank_search_device(templates, w.orbits, [brp.Toplist(100)])


def test_out_of_range_index_flagged(brp, wl):
    w = wl.plain_small()
    # psi = 2.5: the orbital offset falls by ~4.6 samples over the series, so the last
    # samples map past its end.  Keep every sample anyway (steps = unpadded).
    p = brp.parameters(w.n, w.n, 0.05, 3600.0, 2.5, wl.DT)
    assert brp.sample_offset(w.n - 1, p) < -1.0
    p.steps = p.unpadded
    say("[gpu] resample index past the observation is flagged")
    for kw in (dict(packed=True, engine=brp.ENGINE_FUSED), dict(PLAIN_VARIANTS["selected"])):
        with brp.GPU(p, w.x, 256, 0, 1, **kw) as g:
            g.set_template(p)
            g.run()
            with pytest.raises(RuntimeError, match="invalid resample index"):
                g.candidates(0)


@pytest.mark.parametrize("psi", [2.5, 0.7])
def test_out_of_range_indices_in_snapshot(brp, oracle, wl, psi):
    """Source indices past the observation inside blocks whose samples all lie below steps:
    the fused engine's unmasked gather gives way to the masked one for the waves that leave
    the observation, which writes the reference's index rule for every sample
    (src/science.cpp:80) and zero for the samples outside.  psi = 2.5 runs past the end,
    psi = 0.7 stays inside (the unmasked form throughout)."""
    w = wl.plain_small()
    p = brp.parameters(w.n, w.n, 0.05, 3600.0, psi, wl.DT)
    p.steps = p.unpadded
    lib, op = oracle.load(), oracle.as_oparams(p)
    off = np.array([lib.oracle_sample_offset(i, op) for i in range(p.steps)], dtype=np.float32)
    want_idx = np.trunc((np.arange(p.steps, dtype=np.float32) - off).astype(np.float64) + 0.5).astype(np.int32)
    inside = (want_idx >= 0) & (want_idx < p.unpadded)
    assert (not inside.all()) == (psi == 2.5)
    want = np.where(inside, w.x[np.clip(want_idx, 0, p.unpadded - 1)], 0.0).astype(np.float32)
    say(f"[gpu] fused snapshot indices with psi={psi} ({int((~inside).sum())} outside)")
    with brp.GPU(p, w.x, 256, 0, 1, packed=True, engine=brp.ENGINE_FUSED, snapshot=True) as g:
        g.set_template(p)
        g.run()
        res = np.empty(p.n, np.float32)
        idx = np.empty(p.steps, np.int32)
        import ctypes
        rc = brp.load().brp_hip_snapshot(g._h, 0, res.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), None, None, None)
        assert rc == 0
        if psi == 2.5:
            with pytest.raises(RuntimeError, match="invalid resample index"):
                g.candidates(0)
        else:
            g.candidates(0)
    assert np.array_equal(idx, want_idx)
    assert np.array_equal(res[:p.steps], want)


def test_template_dimension_change_rejected(brp, wl):
    """The plain engine keeps the reference's fixed-dimension rule; the fused engine lifts
    it for F and H and for registered series, and still rejects an unregistered length,
    a capacity change and an oversized batch."""
    w = wl.plain_small()
    p = w.templates[0]
    q = p.copy()
    q.fundamental -= 1
    other = brp.parameters(1 << 15, 1 << 15, p.tau, 3600.0, p.psi, p.dt)
    cap = p.copy()
    cap.capacity -= 1
    say("[gpu] set_templates rejects a dimension change")
    with brp.GPU(p, w.x, 256, 0, 2, stockham=True, packed=True, engine=brp.ENGINE_PLAIN) as g:
        with pytest.raises(RuntimeError, match="dimensions changed"):
            g.set_templates([p, q])
    with brp.GPU(p, w.x, 256, 0, 2, packed=True, engine=brp.ENGINE_FUSED) as g:
        g.set_templates([p, q])
        for bad in (other, cap):
            with pytest.raises(RuntimeError, match="dimensions changed"):
                g.set_templates([p, bad])
        with pytest.raises(RuntimeError, match="batch exceeds"):
            g.set_templates([p, p, p])


def test_rpi_rounding_matches_index_rule():
    """The unmasked gather rounds with v_cvt_rpi_i32_f32; it must equal the index rule's
    int(double(x) + 0.5) (src/science.cpp:80) for every float in [-0.5, 2^24]: an exhaustive
    device check of 2.3e9 values (tests/probes/rpi_rounding.hip, built by build())."""
    import pathlib
    import subprocess
    probe = pathlib.Path(__file__).resolve().parent / "probes" / "rpi_rounding"
    if not probe.exists():
        pytest.skip("probe not built (__graft_entry__.build() compiles it)")
    say("[gpu] exhaustive v_cvt_rpi_i32_f32 check")
    proc = subprocess.run([str(probe)], capture_output=True, text=True, timeout=120)
    say(proc.stdout.strip())
    assert proc.returncode == 0 and "0 mismatches" in proc.stdout.splitlines()[-1], proc.stdout + proc.stderr
