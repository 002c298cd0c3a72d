"""Parity checks of a GPU result against the oracle (test infrastructure).

Bars (from the task's parity rules and the reference's own validation,
src/main.cpp:55-80):
  * resample source indices: bit-exact, and the resampled values equal;
  * power: |P_gpu - P_fp64| <= 1e-5 * max(P_fp64[k], mean P) for every bin;
  * harmonic sums of the GPU power: bit-exact against the oracle's harmonic
    summing of that same power (the reference's "isolated" check, here exact);
  * candidates: identical (bin, level) sets except candidates whose power lies
    within 1e-5 relative of their threshold, which are counted and reported.
"""
from __future__ import annotations

import numpy as np

POWER_RTOL = 1e-5
NEAR_THRESHOLD = 1e-5


def power_error(gpu_power: np.ndarray, p64: np.ndarray) -> dict:
    mean = float(np.mean(p64[1:]))
    floor = np.maximum(p64, mean)
    err = np.abs(gpu_power.astype(np.float64) - p64)
    rel = err / floor
    return {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2))),
            "worst_bin": int(rel.argmax()), "ok": bool(rel.max() <= POWER_RTOL)}


def compare_candidates(gpu: np.ndarray, ref: np.ndarray, thresholds) -> dict:
    """Matches (bin, level) sets; differences within NEAR_THRESHOLD of a threshold are allowed."""
    thr = np.asarray(list(thresholds), dtype=np.float64)
    g = {(int(c["bin"]), int(c["level"])): float(c["power"]) for c in gpu}
    r = {(int(c["bin"]), int(c["level"])): float(c["power"]) for c in ref}
    near = 0
    bad = []
    for key in set(g) ^ set(r):
        val = g.get(key, r.get(key))
        if abs(val - thr[key[1]]) <= NEAR_THRESHOLD * thr[key[1]]:
            near += 1
        else:
            bad.append((key, g.get(key), r.get(key)))
    worst = 0.0
    for key in set(g) & set(r):
        worst = max(worst, abs(g[key] - r[key]) / max(abs(r[key]), 1e-30))
    return {"gpu": len(g), "ref": len(r), "mismatches": bad, "near_threshold": near, "max_power_rel": worst}


def _close(a: float, b: float, rtol: float) -> bool:
    return abs(a - b) <= rtol * max(abs(a), abs(b))


def _inversions(order_a: list, pos_b: dict, power: dict, rtol: float) -> tuple[int, list]:
    """Pairs that two orderings of the same keys put the other way round: (near ties, other)."""
    if len(order_a) < 2:
        return 0, []
    pb = np.array([pos_b[k] for k in order_a])
    pw = np.array([power[k] for k in order_a], dtype=np.float64)
    i, j = np.nonzero(np.triu(pb[:, None] > pb[None, :], 1))
    tied = np.abs(pw[i] - pw[j]) <= rtol * np.maximum(np.abs(pw[i]), np.abs(pw[j]))
    bad = [(order_a[a], order_a[b]) for a, b in zip(i[~tied], j[~tied])]
    return int(tied.sum()), bad


def rank_mismatches(gpu: np.ndarray, ref: np.ndarray, top: int = 100, thresholds=None,
                    rtol: float = NEAR_THRESHOLD) -> dict:
    """src/main.cpp:71-79: the top-`top` candidates per level by (power desc, bin asc), bins
    compared rank by rank.  A rank whose bins differ is a near tie when the two candidates
    there have powers within rtol of each other (an order swap), and a candidate that the
    other list lacks lies within rtol of that list's cut (its top-th power when it has
    `top`, else the level threshold); every other difference is a mismatch.
    Returns {"near_ties": n, "mismatches": [(level, rank, gpu_bin, ref_bin), ...]}."""
    near, bad = 0, []
    for h in range(5):
        a = sorted([c for c in gpu if c["level"] == h], key=lambda c: (-float(c["power"]), int(c["bin"])))[:top]
        b = sorted([c for c in ref if c["level"] == h], key=lambda c: (-float(c["power"]), int(c["bin"])))[:top]
        pa = {int(c["bin"]): float(c["power"]) for c in a}
        pb = {int(c["bin"]): float(c["power"]) for c in b}
        thr = float(thresholds[h]) if thresholds is not None else 0.0
        cut_a = float(a[-1]["power"]) if len(a) == top else thr
        cut_b = float(b[-1]["power"]) if len(b) == top else thr
        power = {**pa, **pb}
        for r in range(max(len(a), len(b))):
            ga = int(a[r]["bin"]) if r < len(a) else None
            rb = int(b[r]["bin"]) if r < len(b) else None
            if ga == rb:
                continue
            ok = ga is None or rb is None or _close(power[ga], power[rb], rtol)
            for binv, other, cut in ((ga, pb, cut_b), (rb, pa, cut_a)):
                if binv is not None and binv not in other:
                    ok = ok and cut > 0 and _close(power[binv], cut, rtol)
            if ok:
                near += 1
            else:
                bad.append((h, r, ga, rb))
    return {"near_ties": near, "mismatches": bad}


SATURATED = 320.0  # -log10 Q of a power whose chi-square tail underflows (src/workload.cpp:146-147)


def _row_key(r) -> tuple:
    """A toplist row's identity: its bin and the template (orbit) that produced it."""
    return (int(r.bin), float(r.period), float(r.tau), float(r.psi))


def compare_toplists(gpu_rows, ref_rows, rows: int, thresholds=None, rtol: float = NEAR_THRESHOLD) -> dict:
    """Bank toplist (5 levels x `rows`, src/workload.cpp:126-141) against the oracle loop's.

    Exact equality (the same rows, each from the same template, in the same order, powers
    within rtol) is the expected result.  Differences that 1e-6-level power differences can cause are counted
    as near ties, and every other difference is a mismatch:
      * a row present in one table only, whose power is within rtol of the other table's
        boundary at that level (its last row when full, else the level's base threshold), or
        of a row of the same bin from another template in the other table (the per-bin
        replacement went the other way);
      * two common rows in opposite order whose powers are within rtol of each other;
    and common rows must agree in power within rtol.
    Returns {"exact": bool, "near_ties": n, "mismatches": [...]}."""
    near, bad = 0, []
    exact = True
    for h in range(5):
        g = [r for r in gpu_rows[h * rows:(h + 1) * rows] if r.harmonics]
        f = [r for r in ref_rows[h * rows:(h + 1) * rows] if r.harmonics]
        if [_row_key(r) for r in g] != [_row_key(r) for r in f]:
            exact = False
        base = float(thresholds[h]) if thresholds is not None else 0.0
        bound_g = max(base, float(g[-1].power) if len(g) == rows else 0.0)
        bound_r = max(base, float(f[-1].power) if len(f) == rows else 0.0)
        gk = {_row_key(r): float(r.power) for r in g}
        rk = {_row_key(r): float(r.power) for r in f}
        for k in gk.keys() & rk.keys():
            if not _close(gk[k], rk[k], rtol):
                bad.append(("power", h, k, gk[k], rk[k]))
        only_g, only_r = gk.keys() - rk.keys(), rk.keys() - gk.keys()
        for mine, others, other_pw, bound in ((only_g, only_r, rk, bound_r), (only_r, only_g, gk, bound_g)):
            src = gk if mine is only_g else rk
            for k in mine:
                p = src[k]
                partner = any(j[0] == k[0] and _close(other_pw[j], p, rtol) for j in others)
                if _close(p, bound, rtol) or partner:
                    near += 1
                else:
                    bad.append(("extra", h, k, p, bound))
        order_g = [_row_key(r) for r in g if _row_key(r) in rk]
        pos_r = {k: i for i, k in enumerate(_row_key(r) for r in f if _row_key(r) in gk)}
        tied, inv = _inversions(order_g, pos_r, gk, rtol)
        near += tied
        bad += [("order", h, a, b) for a, b in inv]
    return {"exact": exact and not bad, "near_ties": near, "mismatches": bad}


def compare_finals(gpu_final, ref_final, rtol: float = NEAR_THRESHOLD) -> dict:
    """Toplist::final (src/workload.cpp:142-159) against the oracle's: rows compared by
    (bin, harmonics) in order.  Rows whose significance saturates at 320 in both lists tie
    (final() keeps such rows in the toplist's row order), so they are compared as a set,
    and the saturated rows that sit at different positions are counted.  Outside the
    saturated rows, a position whose rows differ is a near tie when their significances are
    within rtol; anything else is a mismatch.
    Returns {"exact": bool, "saturated": n, "saturated_moved": n, "near_ties": n,
    "mismatches": [...]}."""
    g = [(int(r.bin), int(r.harmonics), float(r.significance)) for r in gpu_final]
    f = [(int(r.bin), int(r.harmonics), float(r.significance)) for r in ref_final]
    sat_g = [(b, hm) for b, hm, s in g if s >= SATURATED]
    sat_f = [(b, hm) for b, hm, s in f if s >= SATURATED]
    bad, near = [], 0
    if sorted(sat_g) != sorted(sat_f):
        bad.append(("saturated set", sorted(set(sat_g) ^ set(sat_f))[:10]))
    moved = sum(a != b for a, b in zip(sat_g, sat_f))
    rest_g = [(b, hm, s) for b, hm, s in g if s < SATURATED]
    rest_f = [(b, hm, s) for b, hm, s in f if s < SATURATED]
    if len(rest_g) != len(rest_f):
        bad.append(("length", len(g), len(f)))
    for i, (a, b) in enumerate(zip(rest_g, rest_f)):
        if a[:2] == b[:2]:
            continue
        if _close(a[2], b[2], rtol):
            near += 1
        else:
            bad.append(("row", i, a, b))
    exact = [x[:2] for x in g] == [x[:2] for x in f]
    return {"exact": exact and not bad, "saturated": len(sat_f), "saturated_moved": moved, "near_ties": near,
            "mismatches": bad}


def check_template(snap, oracle, x, p, label: str = "") -> dict:
    """Full parity of one template's GPU snapshot against the oracle; raises on failure."""
    res, idx, p32, p64, summed, cands = oracle.reference(x, p)
    steps = p.steps
    report = {"case": label}
    if snap.index is not None:
        mism = int(np.count_nonzero(snap.index[:steps] != idx))
        report["index_mismatches"] = mism
        assert mism == 0, f"{label}: {mism} resample index mismatches"
    vmis = int(np.count_nonzero(snap.resampled[:steps] != res[:steps]))
    report["resampled_mismatches"] = vmis
    assert vmis == 0, f"{label}: {vmis} resampled value mismatches"
    pad_err = float(np.max(np.abs(snap.resampled[steps:] - res[steps:]))) if steps < p.n else 0.0
    report["pad_abs_err"] = pad_err
    assert pad_err <= 1e-6 * max(1.0, float(np.abs(res[:steps]).mean())), f"{label}: pad mean differs by {pad_err}"
    pe = power_error(snap.power, p64)
    report["power"] = pe
    assert pe["ok"], f"{label}: power error {pe}"
    iso = oracle.harmonic(snap.power, p)
    diff = int(np.count_nonzero(iso != snap.summed))
    report["harmonic_isolated_mismatches"] = diff
    assert diff == 0, f"{label}: {diff} harmonic sums differ from the oracle on the GPU power"
    cc = compare_candidates(snap.candidates, cands, p.threshold)
    report["candidates"] = {k: v for k, v in cc.items() if k != "mismatches"}
    assert not cc["mismatches"], f"{label}: candidate mismatches {cc['mismatches'][:5]}"
    rk = rank_mismatches(snap.candidates, cands, thresholds=p.threshold)
    report["rank_mismatches"] = len(rk["mismatches"])
    report["rank_near_ties"] = rk["near_ties"]
    assert not rk["mismatches"], f"{label}: top-100 rank mismatches {rk['mismatches'][:5]}"
    return report
