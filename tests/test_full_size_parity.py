"""Oracle parity of the bank workloads at their full sizes (GPU, through the C ABI).

  * the 512-template bank at N = 2^22 (bank semantics), top-100 and top-1000 per level,
    through the host-toplist bank loop and the device-toplist one;
  * the uneven workunit: 4-bit DD blocks at N = 2^20, 3 * 2^19 and 2^22 with pulsars and
    RFI, preprocessed on the device (checked against the oracle's FP64 preprocessing at
    those lengths), then 2048 templates over the three series in one launch schedule,
    top-1000 and top-100 per series.

The reference's bank loop (src/workunit_main.cpp:77-136) is the oracle's one template at
a time: every template's candidates from the FP64-spectrum pipeline on a host thread pool,
then thresholds + retain replayed in template order (tests/oracle_lib.bank_toplists).
Toplists and final lists are compared with parity.compare_toplists / compare_finals:
rows, templates and order must agree, and only differences that 1e-6-level power
differences can cause (rows within 1e-5 of a level's boundary, swaps of powers within
1e-5, significance ties at 320) are allowed; those are counted and printed.
"""
import concurrent.futures
import sys
import time

import numpy as np
import pytest

import parity

pytestmark = [pytest.mark.gpu, pytest.mark.slow, pytest.mark.timeout(400)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_bank(label, gpu_top, ref_top, rows, thresholds):
    rep = parity.compare_toplists(gpu_top.table(), list(ref_top.state), rows, thresholds)
    fin = parity.compare_finals(gpu_top.final(), ref_top.final())
    filled = [sum(1 for r in list(ref_top.state)[h * rows:(h + 1) * rows] if r.harmonics) for h in range(5)]
    say(f"[full] {label}: rows per level {filled}, toplist exact={rep['exact']} near_ties={rep['near_ties']}; "
        f"final exact={fin['exact']} rows={len(ref_top.final())} saturated={fin['saturated']} "
        f"moved={fin['saturated_moved']} near_ties={fin['near_ties']}")
    assert not rep["mismatches"], f"{label}: toplist mismatches {rep['mismatches'][:5]}"
    assert not fin["mismatches"], f"{label}: final mismatches {fin['mismatches'][:5]}"
    assert sum(filled) > 0
    return rep, fin


def test_bank_512_n2p22_vs_oracle_loop(brp, oracle):
    import workloads

    w = workloads.bank(count=512, n=1 << 22)
    n = w.n
    rows_list = (100, 1000)
    say(f"[full] {w.name}: fused bank loop (batch 128, group 8, graph), host and device toplists")
    gpu = {}
    with brp.GPU(w.templates[0], w.x, 256, 0, 128, packed=True, engine=brp.ENGINE_FUSED, graph=True, group=8,
                 untimed=True) as g:
        for rows in rows_list:
            top = brp.Toplist(rows)
            g.bank_search(w.orbits, n, workloads.SAMPLE_US, top)
            gpu[("host", rows)] = top
            say(f"[full] bank_search rows={rows} done")
        for rows in rows_list:
            tops = [brp.Toplist(rows)]
            g.bank_search_device(w.templates, w.orbits, tops)
            gpu[("device", rows)] = tops[0]
            say(f"[full] bank_search_device rows={rows} done")
    t0 = time.time()
    say(f"[full] oracle: 512 templates on {oracle.worker_threads()} host threads")
    cands = oracle.bank_candidates(w.x, w.templates)
    say(f"[full] oracle candidates in {time.time() - t0:.1f} s: {sum(len(c) for c in cands)}")
    refs = dict(zip(rows_list, oracle.bank_toplists(w.templates, w.orbits, cands, rows_list)))
    for (mode, rows), top in gpu.items():
        _check_bank(f"N=2^22 x 512 {mode} toplist rows={rows}", top, refs[rows], rows, w.templates[0].threshold)


def _preprocessed_series(brp, oracle, workloads, lengths):
    """Device preprocessing of the uneven workunit's DD blocks (as the benchmark makes them)
    against the oracle's, at full length: decoded samples bit for bit, the same number of
    RFI bins replaced, rel-L2 <= 1e-5 on the whitened series.  Returns the device series."""
    blocks, xs = [], []
    for k, n in enumerate(lengths):
        block, _ = brp.synthetic_dd(n, workloads.DT, 42 + 10 * k, 7.5, 2.0, 4, workloads.spin_for(42 + 10 * k),
                                    0.05, 0.1, workloads.INJECTED, n, 10, 50.0)
        blocks.append(block)
    say(f"[full] oracle preprocessing of N={list(lengths)} on {len(lengths)} host threads")
    with concurrent.futures.ThreadPoolExecutor(len(lengths)) as pool:
        futs = [pool.submit(oracle.preprocess, b, n, 4, 1.0, n, 1000, 30.0) for b, n in zip(blocks, lengths)]
        for block, n, fut in zip(blocks, lengths, futs):
            say(f"[full] device preprocess N={n}")
            out, dec, rfi = brp.preprocess(block, n, 4, 1.0, n, 1000, 30.0)
            ref, rdec, rrfi = fut.result()
            rel = float(np.linalg.norm(out.astype(np.float64) - ref) / np.linalg.norm(ref))
            say(f"[full] preprocess N={n}: rfi {rfi} (oracle {rrfi}), rel-L2 {rel:.3e}")
            assert np.array_equal(dec, rdec), n
            assert rfi == rrfi and rfi >= 10, n
            assert rel <= 1e-5, n
            xs.append(out)
    return xs


def test_uneven_workunit_2048_full_lengths_vs_oracle_loop(brp, oracle):
    import workloads

    lengths = (1 << 20, 3 << 19, 1 << 22)
    xs = _preprocessed_series(brp, oracle, workloads, lengths)
    u = workloads.uneven(count=2048, lengths=lengths, xs=xs)
    byn = {n: x for n, x, _ in u.series}
    first = brp.search_parameters(lengths[-1], lengths[-1], workloads.SAMPLE_US, workloads.INJECTED)
    reg = [first.n] + [n for n in lengths if n != first.n]
    rows_list = (1000, 100)
    gpu = {}
    say(f"[full] {u.name}: one launch schedule, batch 128, group 8, graph")
    with brp.GPU(first, byn[first.n], 256, 0, 128, packed=True, engine=brp.ENGINE_FUSED, graph=True, group=8,
                 untimed=True) as g:
        for n in reg[1:]:
            g.add_series(brp.search_parameters(n, n, workloads.SAMPLE_US, workloads.INJECTED), byn[n])
        for rows in rows_list:
            tops = [brp.Toplist(rows) for _ in reg]
            counts, _ = g.bank_search_templates(u.templates, u.orbits, tops)
            assert counts.sum() > 0
            gpu[rows] = dict(zip(reg, tops))
            say(f"[full] bank_search_templates rows={rows} done")
    for n in reg:
        sel = [i for i, p in enumerate(u.templates) if p.n == n]
        tpl = [u.templates[i] for i in sel]
        orb = [u.orbits[i] for i in sel]
        t0 = time.time()
        cands = oracle.bank_candidates(byn[n], tpl)
        say(f"[full] oracle N={n}: {len(sel)} templates in {time.time() - t0:.1f} s")
        refs = dict(zip(rows_list, oracle.bank_toplists(tpl, orb, cands, rows_list)))
        for rows in rows_list:
            _check_bank(f"uneven N={n} ({len(sel)} templates) rows={rows}", gpu[rows][n], refs[rows], rows,
                        tpl[0].threshold)
