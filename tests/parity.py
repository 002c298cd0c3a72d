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


def rank_mismatches(gpu: np.ndarray, ref: np.ndarray, top: int = 100) -> int:
    """src/main.cpp:71-79: top-100 per level by (power desc, bin asc), bins compared by rank."""
    bad = 0
    for h in range(5):
        a = sorted([c for c in gpu if c["level"] == h], key=lambda c: (-float(c["power"]), int(c["bin"])))
        b = sorted([c for c in ref if c["level"] == h], key=lambda c: (-float(c["power"]), int(c["bin"])))
        for i in range(min(top, len(a), len(b))):
            bad += int(a[i]["bin"] != b[i]["bin"])
    return bad


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
    report["rank_mismatches"] = rank_mismatches(snap.candidates, cands)
    return report
