"""The toplist and rank comparison rules of tests/parity.py (CPU).

Each rule is checked both ways: a difference that 1e-6-level power differences can cause
is counted as a near tie, and a plausible real mistake (a missing strong row, a swap of
well-separated powers, a wrong power, a different final row) is reported as a mismatch.
"""
from dataclasses import dataclass

import numpy as np

import parity


@dataclass
class Row:
    power: float = 0.0
    period: float = 0.0
    tau: float = 0.0
    psi: float = 0.0
    significance: float = 0.0
    harmonics: int = 0
    bin: int = 0


def _table(levels, rows):
    """levels: per level a list of (bin, power, template id); padded with empty rows."""
    out = []
    for h, lv in enumerate(levels):
        lv = sorted(lv, key=lambda e: -e[1])
        for b, p, t in lv:
            out.append(Row(power=p, period=1000.0 + t, tau=0.1, psi=float(t), harmonics=1 << h, bin=b))
        out += [Row() for _ in range(rows - len(lv))]
    return out


BASE = [[(600 + i, 40.0 + 3 * i, i) for i in range(5)] for _ in range(5)]
THR = [17.0, 20.0, 25.0, 33.0, 47.0]


def _levels(**edits):
    lv = [list(x) for x in BASE]
    for h, entries in edits.items():
        lv[int(h[1:])] = entries
    return lv


def test_identical_tables_are_exact():
    rep = parity.compare_toplists(_table(BASE, 5), _table(BASE, 5), 5, THR)
    assert rep == {"exact": True, "near_ties": 0, "mismatches": []}


def test_swap_of_near_equal_powers_is_a_near_tie():
    a = _levels(h2=[(600, 50.0, 0), (601, 50.0000001, 1), (602, 44.0, 2), (603, 43.0, 3), (604, 42.0, 4)])
    b = _levels(h2=[(600, 50.0000002, 0), (601, 50.0, 1), (602, 44.0, 2), (603, 43.0, 3), (604, 42.0, 4)])
    rep = parity.compare_toplists(_table(a, 5), _table(b, 5), 5, THR)
    assert not rep["mismatches"] and rep["near_ties"] == 1 and not rep["exact"]


def test_swap_of_separated_powers_is_a_mismatch():
    a = _levels(h1=[(600, 50.0, 0), (601, 49.0, 1)])
    b = _levels(h1=[(600, 49.0, 0), (601, 50.0, 1)])  # same rows, powers (and so order) exchanged
    rep = parity.compare_toplists(_table(a, 5), _table(b, 5), 5, THR)
    assert rep["mismatches"]


def test_power_disagreement_is_a_mismatch():
    a = _levels(h0=[(600, 50.0, 0)])
    b = _levels(h0=[(600, 50.01, 0)])
    rep = parity.compare_toplists(_table(a, 5), _table(b, 5), 5, THR)
    assert [m[0] for m in rep["mismatches"]] == ["power"]


def test_row_at_the_boundary_is_a_near_tie():
    # full tables (3 rows): the last row differs, both within 1e-5 of the other's boundary
    a = _levels(h3=[(600, 90.0, 0), (601, 80.0, 1), (602, 70.0000001, 2)])
    b = _levels(h3=[(600, 90.0, 0), (601, 80.0, 1), (777, 70.0000004, 9)])
    for lv in (a, b):
        for h in range(5):
            lv[h] = lv[h][:3]
    rep = parity.compare_toplists(_table(a, 3), _table(b, 3), 3, THR)
    assert not rep["mismatches"] and rep["near_ties"] == 2


def test_missing_strong_row_is_a_mismatch():
    a = _levels(h4=[(600, 300.0, 0), (601, 60.0, 1)])
    b = _levels(h4=[(601, 60.0, 1)])  # not full: its boundary is the threshold 47
    rep = parity.compare_toplists(_table(a, 5), _table(b, 5), 5, THR)
    assert [m[0] for m in rep["mismatches"]] == ["extra"]


def test_row_near_the_threshold_of_a_partial_table_is_a_near_tie():
    a = _levels(h4=[(600, 300.0, 0), (601, 47.0000002, 1)])
    b = _levels(h4=[(600, 300.0, 0)])
    rep = parity.compare_toplists(_table(a, 5), _table(b, 5), 5, THR)
    assert not rep["mismatches"] and rep["near_ties"] == 1


def test_per_bin_replacement_tie_is_a_near_tie():
    # the same bin kept from two different templates whose powers are within 1e-5
    a = _levels(h0=[(650, 80.0000003, 3)])
    b = _levels(h0=[(650, 80.0, 4)])
    rep = parity.compare_toplists(_table(a, 5), _table(b, 5), 5, THR)
    assert not rep["mismatches"] and rep["near_ties"] == 2


def _final(entries):
    return [Row(power=1.0, significance=s, harmonics=hm, bin=b) for b, hm, s in entries]


def test_finals_saturated_rows_compare_as_a_set():
    a = _final([(10, 1, 320.0), (11, 2, 320.0), (12, 1, 55.0), (13, 4, 40.0)])
    b = _final([(11, 2, 320.0), (10, 1, 320.0), (12, 1, 55.0), (13, 4, 40.0)])
    rep = parity.compare_finals(a, b)
    assert not rep["mismatches"] and rep["saturated"] == 2 and rep["saturated_moved"] == 2
    c = _final([(11, 2, 320.0), (99, 1, 320.0), (12, 1, 55.0), (13, 4, 40.0)])
    assert parity.compare_finals(a, c)["mismatches"]


def test_finals_near_and_real_differences():
    a = _final([(12, 1, 55.0), (13, 4, 55.0000001), (14, 1, 30.0)])
    b = _final([(13, 4, 55.0000001), (12, 1, 55.0), (14, 1, 30.0)])
    rep = parity.compare_finals(a, b)
    assert not rep["mismatches"] and rep["near_ties"] == 2
    c = _final([(12, 1, 55.0), (13, 4, 55.0000001), (15, 1, 29.0)])
    assert parity.compare_finals(a, c)["mismatches"]
    assert parity.compare_finals(a, a[:2])["mismatches"]  # a lost row


def _cands(entries):
    return np.array([(b, lv, p, 0) for b, lv, p in entries], dtype=parity_dtype())


def parity_dtype():
    return np.dtype([("bin", "<u4"), ("level", "<u4"), ("power", "<f4"), ("reserved", "<u4")])


def test_rank_rule():
    base = [(600 + i, 0, 100.0 - i) for i in range(10)]
    assert parity.rank_mismatches(_cands(base), _cands(base), top=5, thresholds=THR) == \
        {"near_ties": 0, "mismatches": []}
    a, b = list(base), list(base)
    a[2], a[3] = (602, 0, 97.00001), (603, 0, 97.0)  # float32 powers 8e-6 apart, opposite order
    b[2], b[3] = (602, 0, 97.0), (603, 0, 97.00001)
    assert parity.rank_mismatches(_cands(a), _cands(b), top=5, thresholds=THR) == \
        {"near_ties": 2, "mismatches": []}
    a[2], a[3] = (602, 0, 97.1), (603, 0, 97.0)  # 1e-3 apart: a real order difference
    b[2], b[3] = (602, 0, 97.0), (603, 0, 97.1)
    assert len(parity.rank_mismatches(_cands(a), _cands(b), top=5, thresholds=THR)["mismatches"]) == 2
    far = list(base)
    far[1] = (601, 0, 50.0)  # ranks 1..4 shift: a real difference
    assert parity.rank_mismatches(_cands(far), _cands(base), top=5, thresholds=THR)["mismatches"]


def test_fp32_and_fp64_oracle_banks_agree_under_the_rules(oracle):
    """The oracle's FP32 FFT and power against its FP64 ones (per-bin differences near 1e-6,
    like the GPU's) over a 64-template bank at N = 2^20: the bank toplists and final lists
    meet the comparison rules for 100 and 1000 rows."""
    import workloads

    w = workloads.bank(count=64, n=1 << 20)
    c64 = oracle.bank_candidates(w.x, w.templates)
    c32 = oracle.bank_candidates(w.x, w.templates, fft32=True)
    for rows in (100, 1000):
        t64, = oracle.bank_toplists(w.templates, w.orbits, c64, (rows,))
        t32, = oracle.bank_toplists(w.templates, w.orbits, c32, (rows,))
        rep = parity.compare_toplists(list(t32.state), list(t64.state), rows, w.templates[0].threshold)
        fin = parity.compare_finals(t32.final(), t64.final())
        assert not rep["mismatches"] and not fin["mismatches"], rows
        assert sum(r.harmonics > 0 for r in t64.state) > 0
