"""Synthetic sector workloads for batching and serving.

The reference runs real cyclic sectors, one process each
(scripts/reproduce.py:235-276); it defines no synthetic mixes.  The two mixes
below are drawn with the reference's own splitmix64 stream and
rejection-bounded draws (cuda/cuprym_cyclic.cu:1153-1158, 1207-1218), so they
are reproducible bit for bit from their seeds:

* `uneven_sectors(seed=3, count=64)`: n in [12,22], k in [3,6], M in
  [20,300], M' = M + ceil(M/10), p = 65521, b in {8,16}; a draw whose
  N = C(n,k) M falls outside [1e3, 1e7] is redrawn.
* `poisson_stream(seed=5, count, rate=2.0)`: arrival gaps exponential with
  rate 2/s; N log-uniform in [1e4, 3e7]; (p, b) uniform over
  {65521, 2^31-1} x {8, 16}; n in [12,24], k in [3,6] redrawn until
  M = round(N / C(n,k)) lies in [8, 400], M' = M + ceil(M/10).

Each sector i of a mix gets tables from seeds (1000 + 2i, 1001 + 2i), its
Dr, Dc and V from the reference recipe (pgw_reference_random_state with the
mix seed and sector i) and two row-mixing rounds from seed 10001.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from . import SyntheticSpec

MASK = (1 << 64) - 1


class SplitMix:
    """splitmix64 stream with the reference's bounded draws."""

    def __init__(self, seed: int):
        self.state = seed & MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def bounded(self, bound: int) -> int:
        threshold = ((1 << 64) - bound) % bound
        while True:
            v = self.next()
            if v >= threshold:
                return v % bound

    def uniform(self) -> float:
        """(0, 1] from the top 53 bits."""
        return ((self.next() >> 11) + 1) / float(1 << 53)


@dataclass
class SectorJob:
    index: int
    spec: SyntheticSpec
    block: int
    arrival: float = 0.0  # seconds after the stream starts (Poisson mixes)

    @property
    def cols(self) -> int:
        return self.spec.cols

    def full_length(self) -> int:
        """The reference's target length 2N/b + 8 for the synthetic configs."""
        return 2 * self.cols // self.block + 8


def _mprime(m: int) -> int:
    return m + (m + 9) // 10


def uneven_sectors(seed: int = 3, count: int = 64, prime: int = 65521,
                   n_lo: int = 12, n_hi: int = 22, k_lo: int = 3, k_hi: int = 6,
                   m_lo: int = 20, m_hi: int = 300, n_min: int = 1000,
                   n_max: int = 10_000_000) -> list[SectorJob]:
    rng = SplitMix(seed)
    out = []
    while len(out) < count:
        n = n_lo + rng.bounded(n_hi - n_lo + 1)
        k = k_lo + rng.bounded(k_hi - k_lo + 1)
        m = m_lo + rng.bounded(m_hi - m_lo + 1)
        b = (8, 16)[rng.bounded(2)]
        N = math.comb(n, k) * m
        if not (n_min <= N <= n_max):
            continue
        i = len(out)
        out.append(SectorJob(i, SyntheticSpec(n, k, m, _mprime(m), prime, 1000 + 2 * i,
                                              1001 + 2 * i), b))
    return out


def poisson_stream(seed: int = 5, count: int = 32, rate: float = 2.0,
                   n_min: float = 1e4, n_max: float = 3e7) -> list[SectorJob]:
    rng = SplitMix(seed)
    out = []
    t = 0.0
    for i in range(count):
        t += -math.log(rng.uniform()) / rate
        target = math.exp(math.log(n_min) + (math.log(n_max) - math.log(n_min)) * rng.uniform())
        p, b = ((65521, 8), (65521, 16), (2147483647, 8), (2147483647, 16))[rng.bounded(4)]
        while True:
            n = 12 + rng.bounded(13)
            k = 3 + rng.bounded(4)
            m = round(target / math.comb(n, k))
            if 8 <= m <= 400:
                break
        out.append(SectorJob(i, SyntheticSpec(n, k, m, _mprime(m), p, 1000 + 2 * i, 1001 + 2 * i),
                             b, arrival=t))
    return out


def sector_state(job: SectorJob, mix_seed: int, rounds: int = 2, rowmix_seed: int = 10001):
    """Dr, Dc, V and the row-mixing plan of one sector of a mix."""
    from . import reference_random_state, rowmix_plan, rowmix_realized_seed

    N, p = job.cols, job.spec.prime
    dr, dc, v = reference_random_state(mix_seed, job.index, p, N, N, job.block)
    realized = rowmix_realized_seed(rowmix_seed, "synthetic", job.index, N, p)
    pairs = rowmix_plan(realized, N, rounds, p) if rounds else None
    return dr, dc, v, pairs
