"""The measured workloads, shared by the bench and the full-size GPU tests so
both draw the same job sets.

batched_set_sizes: 256 BN254 G1 MSMs of floor(2^U) points, U uniform on
    [10, 18], over one shared SRS prefix, plus 256 BN254-Fr NTTs of 2^K
    elements, K uniform on {12, ..., 20}.
serving mix: 70% BLS12-381 commitments to 4096 evaluations, 30% BN254
    commitments to 2^K evaluations, K uniform on {14, ..., 20}.
"""
from __future__ import annotations

import random

import numpy as np

BATCH_SRS_POINTS = 1 << 18
SERVE_BN254_LOG = (14, 20)
SERVE_BLS_N = 4096


def batched_set_sizes(seed: int = 4242, count: int = 256):
    """(msm sizes, NTT log2 sizes) of the batched job set."""
    rnd = random.Random(seed)
    msm_n = [int(2 ** rnd.uniform(10, 18)) for _ in range(count)]
    ntt_log = [rnd.randint(12, 20) for _ in range(count)]
    return msm_n, ntt_log


def ntt_layout(ntt_log):
    """Element offsets of the NTT jobs laid out by size (jobs of one size are
    contiguous, so each size group is one batched launch sequence), and the
    groups {log2 n: [job, ...]} in ascending size."""
    order = sorted(range(len(ntt_log)), key=lambda j: ntt_log[j])
    off_of, off = {}, 0
    for j in order:
        off_of[j] = off
        off += 1 << ntt_log[j]
    groups = {}
    for j in order:
        groups.setdefault(ntt_log[j], []).append(j)
    return off_of, off, dict(sorted(groups.items()))


def serving_request(rng: np.random.Generator, pool: dict, distinct: bool = False):
    """One request of the serving mix: (curve, evaluations as a uint8 array).
    pool[curve] holds uniform canonical evaluations (pool[2] at least 4096,
    pool[1] at least 2^20 elements).  distinct=True takes each request's
    window at a random element offset into its pool (pools then need room
    past the largest size), so requests of one size differ."""
    if rng.random() < 0.7:
        curve, n = 2, SERVE_BLS_N
    else:
        curve, n = 1, 1 << int(rng.integers(SERVE_BN254_LOG[0], SERVE_BN254_LOG[1] + 1))
    room = pool[curve].size // 32 - n
    off = int(rng.integers(0, room + 1)) if distinct and room > 0 else 0
    return curve, pool[curve][32 * off: 32 * (off + n)]
