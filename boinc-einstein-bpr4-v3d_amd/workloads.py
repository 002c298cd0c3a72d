"""Seeded synthetic workloads of the BRP4 per-template search.

Each builder returns the observation (float32 series) and the template parameter
blocks, made with the library's own generator (counter-based RNG, Box-Muller in
double) so that every run and every machine sees the same bytes.

  plain_small   N = 65536, one matching template, bench semantics
  plain_full    N = 2^22, one matching template, random spin in [50, 500] Hz
  bank          N = 2^22 (or smaller), seeded orbit bank, bank semantics
  mixed_radix   N = 3 * 2^20, bench semantics, several templates
  uneven        several series lengths (2^k and 3 * 2^k), one seeded bank split over them
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

import brp_hip as brp

DT = 64e-6
SAMPLE_US = 64.0
INJECTED = (3600.0, 0.05, 1.0)  # P_orb [s], tau [s], psi [rad]


@dataclass
class Workload:
    name: str
    x: np.ndarray
    templates: list
    orbits: list = field(default_factory=list)
    n: int = 0
    spin_hz: float = 0.0


def _series(n: int, seed: int, spin_hz: float, amplitude: float = 0.05, orbit=INJECTED) -> np.ndarray:
    return brp.synthetic_series(n, DT, seed, 0.0, 1.0, spin_hz, 0.05, amplitude, orbit)


def plain_small(seed: int = 42, threshold_scale: float = 1.0, flags: int = 0) -> Workload:
    """N = 65536 noise + 200 Hz train (duty 5%, 0.05 sigma) in a circular orbit, matching template."""
    n = 65536
    x = _series(n, seed, 200.0)
    p = brp.parameters(n, n, INJECTED[1], INJECTED[0], INJECTED[2], DT)
    if flags:
        p.flags = flags
        p.s0 = float(np.float32(np.float32(p.tau) * np.float32(math.sin(np.float32(p.psi)))) * np.float32(p.step_inv))
        brp.resampling_steps(p)
    for h in range(5):
        p.threshold[h] *= threshold_scale
    return Workload("plain N=65536", x, [p], [INJECTED], n, 200.0)


def spin_for(seed: int) -> float:
    return 50.0 + 450.0 * brp.uniform(seed, 1 << 40)


def plain_full(seed: int = 42, n: int = 1 << 22, threshold_scale: float = 1.0) -> Workload:
    """N = 2^22 noise + train with a seeded spin frequency in [50, 500] Hz, matching template."""
    spin = spin_for(seed)
    x = _series(n, seed, spin)
    p = brp.parameters(n, n, INJECTED[1], INJECTED[0], INJECTED[2], DT)
    for h in range(5):
        p.threshold[h] *= threshold_scale
    return Workload(f"plain N={n}", x, [p], [INJECTED], n, spin)


def bank(count: int = 512, seed: int = 42, n: int = 1 << 22, include_injected: bool = True) -> Workload:
    """The plain_full series against `count` seeded orbits (P log-uniform [660, 10000] s,
    tau uniform [0, 0.3] s, psi uniform [0, 2 pi)); bank semantics (f0 400 Hz, false
    alarm 0.08, window 1000).  The injected orbit is template 0 when include_injected."""
    spin = spin_for(seed)
    x = _series(n, seed, spin)
    orbits = [(o.period, o.tau, o.psi) for o in brp.synthetic_bank(seed + 1, count)]
    if include_injected:
        orbits[0] = INJECTED
    templates = [brp.search_parameters(n, n, SAMPLE_US, o) for o in orbits]
    return Workload(f"bank {count} x N={n}", x, templates, orbits, n, spin)


def mixed_radix(count: int = 4, seed: int = 42, n: int = 3 << 20) -> Workload:
    """N = 3 * 2^20, bench semantics, the injected template plus seeded orbits."""
    spin = spin_for(seed)
    x = _series(n, seed, spin)
    orbits = [INJECTED] + [(o.period, o.tau, o.psi) for o in brp.synthetic_bank(seed + 2, count - 1)]
    templates = [brp.parameters(n, n, o[1], o[0], o[2], DT) for o in orbits]
    return Workload(f"mixed radix {count} x N={n}", x, templates, orbits, n, spin)


def algorithmic_bytes(n: int, unpadded: int, harmonic: int) -> int:
    """Compulsory HBM traffic of one template in a fused pipeline: read the series once
    (4 U), one four-step exchange of N/2 complex64 values written and read back (8 N),
    the power below H written and read back (8 H)."""
    return 4 * unpadded + 8 * n + 8 * harmonic


@dataclass
class Uneven:
    name: str
    series: list            # [(n, x float32, spin_hz)] in registration order
    templates: list         # bank-semantics parameter blocks, any series
    orbits: list
    series_of: list         # series index of each template


def uneven(count: int = 2048, seed: int = 42, lengths=(1 << 20, 3 << 19, 1 << 22), xs=None) -> Uneven:
    """One observation per length (seed + 10 k, its own spin frequency, the injected orbit)
    and `count` seeded orbits split over the lengths by a seeded uniform draw; the first
    template of each length is the injected orbit.  Bank semantics on each series.  xs
    (optional) replaces the generated float series (for example by preprocessed ones)."""
    series = []
    for k, n in enumerate(lengths):
        spin = spin_for(seed + 10 * k)
        x = _series(n, seed + 10 * k, spin) if xs is None else np.ascontiguousarray(xs[k], np.float32)
        series.append((n, x, spin))
    orbits = [(o.period, o.tau, o.psi) for o in brp.synthetic_bank(seed + 1, count)]
    which = [min(int(len(lengths) * brp.uniform(seed + 7, i)), len(lengths) - 1) for i in range(count)]
    seen = set()
    for i, s in enumerate(which):
        if s not in seen:
            orbits[i] = INJECTED
            seen.add(s)
    templates = [brp.search_parameters(lengths[s], lengths[s], SAMPLE_US, o) for s, o in zip(which, orbits)]
    return Uneven(f"uneven {count} over N={list(lengths)}", series, templates, orbits, which)

