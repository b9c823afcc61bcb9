"""Textbook finite-field and short-Weierstrass arithmetic on Python integers.

Independent of the oracle's Montgomery/Jacobian code: affine chord-and-tangent
formulas with Fermat inversion via pow().  Used by the tests to pin the oracle
and the HIP path on small cases (it is far too slow for large ones).
"""
from __future__ import annotations

BLS_P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
BLS_R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
BN_P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
BN_R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
SECP_P = 2**256 - 2**32 - 977
SECP_N = 0xFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEBAAEDCE6AF48A03BBFD25E8CD0364141

CURVES = {
    0: dict(p=SECP_P, r=SECP_N, b=7, fb=32,
            g=(0x79BE667EF9DCBBAC55A06295CE870B07029BFCDB2DCE28D959F2815B16F81798,
               0x483ADA7726A3C4655DA4FBFC0E1108A8FD17B448A68554199C47D08FFB10D4B8)),
    1: dict(p=BN_P, r=BN_R, b=3, fb=32, g=(1, 2)),
    2: dict(p=BLS_P, r=BLS_R, b=4, fb=48,
            g=(0x17F1D3A73197D7942695638C4FA9AC0FC3688C4F9774B905A14E3A3F171BAC586C55E83FF97A1AEFFB3AF00ADB22C6BB,
               0x08B3F481E3AAA0F1A09E30ED741D8AE4FCF5E095D5D00AF600DB18CB2C04B3EDD03CC744A2888AE40CAA232946C5E7E1)),
}

INF = None


def on_curve(curve, P):
    if P is INF:
        return True
    c = CURVES[curve]
    x, y = P
    return (y * y - x * x * x - c["b"]) % c["p"] == 0


def add(curve, P, Q):
    p = CURVES[curve]["p"]
    if P is INF:
        return Q
    if Q is INF:
        return P
    x1, y1 = P
    x2, y2 = Q
    if x1 == x2:
        if (y1 + y2) % p == 0:
            return INF
        lam = 3 * x1 * x1 * pow(2 * y1, p - 2, p) % p
    else:
        lam = (y2 - y1) * pow(x2 - x1, p - 2, p) % p
    x3 = (lam * lam - x1 - x2) % p
    return (x3, (lam * (x1 - x3) - y1) % p)


def neg(curve, P):
    if P is INF:
        return INF
    return (P[0], (-P[1]) % CURVES[curve]["p"])


def mul(curve, P, k):
    R = INF
    A = P
    while k > 0:
        if k & 1:
            R = add(curve, R, A)
        A = add(curve, A, A)
        k >>= 1
    return R


def to_wire(curve, P):
    fb = CURVES[curve]["fb"]
    if P is INF:
        return bytes(2 * fb)
    return P[0].to_bytes(fb, "big") + P[1].to_bytes(fb, "big")


def from_wire(curve, w):
    fb = CURVES[curve]["fb"]
    if not any(w):
        return INF
    return (int.from_bytes(w[:fb], "big"), int.from_bytes(w[fb:2 * fb], "big"))


def sqrt(curve, a):
    """sqrt for p = 3 mod 4 (both base fields)."""
    p = CURVES[curve]["p"]
    s = pow(a, (p + 1) // 4, p)
    return s if s * s % p == a % p else None


def compress(curve, P):
    c = CURVES[curve]
    fb, p = c["fb"], c["p"]
    if P is INF:
        out = bytearray(fb)
        out[0] = 0xC0 if curve == 2 else 0x40
        return bytes(out)
    x, y = P
    out = bytearray(x.to_bytes(fb, "big"))
    larger = y > p - y
    if curve == 2:
        out[0] |= 0x80 | (0x20 if larger else 0)
    else:
        out[0] |= 0xC0 if larger else 0x80
    return bytes(out)


def dft(vals, w, r):
    n = len(vals)
    return [sum(vals[j] * pow(w, j * k, r) for j in range(n)) % r for k in range(n)]
