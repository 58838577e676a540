"""Synthetic workloads of the AdaBool OR-filtering path.

Terminology: a *key* (term) is the query unit and inverted-list id; a *value*
(row id) is a bit of the result.  Each workload is seeded: master seed
0xADAB0000 + workload number, with the key-rank permutation shared by the
filter and its queries.

C1  Plain method, small: 100k values x 10k keys, row lengths U{5..15},
    Zipf s=1.0; one query of 64 distinct uniform keys.
C2  Plain method, full size: 4M values x 1M keys, row lengths
    round(LogNormal(ln 16, 0.668)) in [1, 256] (~80M pairs), Zipf s=1.1;
    1024 queries with LogNormal(ln 200, 1.2) sizes in [8, 20000], keys drawn
    Zipf-weighted without replacement.
C3  Batched: the C2 filter; 65,536 queries with power-law sizes
    P(n) ~ n^-1.5 on [1, 50000], Zipf-weighted keys.
C5  Serving: 16M values x 4M keys, LogNormal(ln 20, 0.668) rows in [1, 256]
    (~400M pairs), Zipf s=1.1; query mix of C3.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Tuple

import numpy as np


@dataclass(frozen=True)
class FilterSpec:
    name: str
    seed: int
    nvalues: int
    nkeys: int
    len_dist: int  # 0 uniform [lo, hi]; 1 lognormal(ln p1, p2) clipped
    p1: float
    p2: float
    lo: int
    hi: int
    zipf_s: float


@dataclass(frozen=True)
class QuerySpec:
    seed: int
    nq: int
    size_dist: int  # 0 fixed lo; 1 lognormal(ln p1, p2); 2 power law exponent p1
    p1: float
    p2: float
    lo: int
    hi: int
    key_dist: int  # 0 uniform, 1 Zipf-weighted
    zipf_s: float


C1 = FilterSpec("C1", 0xADAB0001, 100_000, 10_000, 0, 0.0, 0.0, 5, 15, 1.0)
C1_Q = QuerySpec(0xADAB0001, 1, 0, 0.0, 0.0, 64, 64, 0, 1.0)
C2 = FilterSpec("C2", 0xADAB0002, 4_000_000, 1_000_000, 1, 16.0, 0.668, 1, 256, 1.1)
C2_Q = QuerySpec(0xADAB0002, 1024, 1, 200.0, 1.2, 8, 20_000, 1, 1.1)
C3_Q = QuerySpec(0xADAB0003, 65_536, 2, 1.5, 0.0, 1, 50_000, 1, 1.1)
C5 = FilterSpec("C5", 0xADAB0005, 16_000_000, 4_000_000, 1, 20.0, 0.668, 1, 256, 1.1)


def scaled(spec: FilterSpec, factor: float) -> FilterSpec:
    """The same distributions over factor x the values and keys."""
    return replace(spec, name=f"{spec.name}x{factor:g}", nvalues=max(256, int(spec.nvalues * factor)),
                   nkeys=max(16, int(spec.nkeys * factor)))


def make_filter(spec: FilterSpec, nthreads: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    from . import gen_pairs
    return gen_pairs(spec.seed, spec.nvalues, spec.nkeys, len_dist=spec.len_dist, p1=spec.p1, p2=spec.p2,
                     lo=spec.lo, hi=spec.hi, zipf_s=spec.zipf_s, perm_seed=spec.seed, nthreads=nthreads)


def make_queries(fspec: FilterSpec, q: QuerySpec, nq: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    from . import gen_queries
    return gen_queries(q.seed, fspec.nkeys, nq or q.nq, size_dist=q.size_dist, p1=q.p1, p2=q.p2, lo=q.lo,
                       hi=q.hi, key_dist=q.key_dist, zipf_s=q.zipf_s, perm_seed=fspec.seed)
