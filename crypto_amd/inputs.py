"""Seeded synthetic inputs: Fr scalars uniform in [0, r), SRS exponents.

Generator: SplitMix64 (Steele, Lea, Flood 2014); candidate = 4 consecutive
64-bit outputs as little-endian limbs, top limb masked to the bit length of r,
accepted when below r (rejection sampling, so the accepted values are uniform
in [0, r)).  Vectorised with numpy; the CPU tests check it draws exactly the
same values as the C generator in the test oracle.
"""
from __future__ import annotations

import numpy as np

BLS12_381_R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
BN254_R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
SECP256K1_N = 0xFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEBAAEDCE6AF48A03BBFD25E8CD0364141
BANDERWAGON_R = 0x1CFB69D4CA675F520CCE760202687600FF8F87007419047174FD06B52876E7E1
ORDER = {0: SECP256K1_N, 1: BN254_R, 2: BLS12_381_R, 3: BANDERWAGON_R}

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs number start .. start+count-1 of SplitMix64 seeded with `seed`."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
    return z ^ (z >> np.uint64(31))


def scalars(curve: int, seed: int, n: int) -> np.ndarray:
    """n scalars uniform in [0, r) as an (n, 4) uint64 array of LE limbs."""
    r = ORDER[curve]
    rl = [np.uint64((r >> (64 * i)) & 0xFFFFFFFFFFFFFFFF) for i in range(4)]
    top_bits = r.bit_length() - 192
    mask = np.uint64((1 << top_bits) - 1)
    out = np.empty((0, 4), dtype=np.uint64)
    drawn = 0
    while out.shape[0] < n:
        want = n - out.shape[0]
        batch = int(want * 1.4) + 64
        c = splitmix64(seed, drawn * 4, batch * 4).reshape(batch, 4)
        drawn += batch
        c[:, 3] &= mask
        l0, l1, l2, l3 = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
        lt = (l3 < rl[3]) | ((l3 == rl[3]) & ((l2 < rl[2]) | ((l2 == rl[2]) & (
            (l1 < rl[1]) | ((l1 == rl[1]) & (l0 < rl[0]))))))
        out = np.concatenate([out, c[lt]])
    return out[:n]


def scalars_bytes(curve: int, seed: int, n: int) -> bytes:
    """Same scalars as 32-byte little-endian records."""
    return scalars(curve, seed, n).astype("<u8").tobytes()


def to_int(limbs: np.ndarray) -> list[int]:
    """(n, 4) LE uint64 limbs -> Python ints."""
    return [int(a) | (int(b) << 64) | (int(c) << 128) | (int(d) << 192) for a, b, c, d in limbs.tolist()]


def dot_mod(curve: int, a: np.ndarray, b: np.ndarray) -> int:
    """sum a_i * b_i mod r over two (n, 4) limb arrays (closed-form MSM check)."""
    r = ORDER[curve]
    acc = 0
    for x, y in zip(to_int(a), to_int(b)):
        acc += x * y
    return acc % r


def powers(curve: int, tau: int, n: int) -> np.ndarray:
    """tau^i mod r for i < n as (n, 4) LE limbs."""
    r = ORDER[curve]
    out = np.empty((n, 4), dtype=np.uint64)
    v = 1
    m = (1 << 64) - 1
    for i in range(n):
        out[i] = (v & m, (v >> 64) & m, (v >> 128) & m, v >> 192)
        v = v * tau % r
    return out
