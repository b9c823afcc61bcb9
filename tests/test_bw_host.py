"""Banderwagon point operations of the library (crypto_amd/csrc/curve_ed.h:
extended twisted Edwards, unified add-2008-hwcd / dbl-2008-hwcd, their mixed
and quad-cooperative forms) compiled for the host, against the oracle's
projective restatement of the reference's add-2008-bbjlp / dbl-2008-bbjlp.

Inputs carry random projective scales, and cover the identity, the two-torsion
point (0, -1), a point with its negation, equal points (the unified addition
doubling) and the other representative (-x, -y) of a point."""
import ctypes
import os
import random
import subprocess

import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "crypto_amd", "lib", "libkinet_hostcheck.so")
Q = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
R_BW = 0x1CFB69D4CA675F520CCE760202687600FF8F87007419047174FD06B52876E7E1
IDENTITY = bytes(32) + (1).to_bytes(32, "big")
TWO_TORSION = bytes(32) + (Q - 1).to_bytes(32, "big")


@pytest.fixture(scope="module")
def hc():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "hostcheck")], check=True)
    return ctypes.CDLL(LIB)


def neg(w):
    return ((Q - int.from_bytes(w[:32], "big")) % Q).to_bytes(32, "big") + w[32:]


def other_rep(w):
    x, y = int.from_bytes(w[:32], "big"), int.from_bytes(w[32:], "big")
    return ((Q - x) % Q).to_bytes(32, "big") + ((Q - y) % Q).to_bytes(32, "big")


def op(hc, code, a, b, rnd):
    za = rnd.randrange(1, Q).to_bytes(32, "little")
    zb = rnd.randrange(1, Q).to_bytes(32, "little")
    out = ctypes.create_string_buffer(64)
    rc = hc.hc_bw_op(code, a, b, za, zb, out)
    assert rc == 1, f"op {code}: rc {rc}"
    return out.raw


def pairs(rnd):
    g = O.bw_generator()
    pts = [O.bw_mul_gen(rnd.randrange(1, R_BW)) for _ in range(6)]
    out = [(pts[0], pts[1]), (pts[2], pts[3]), (pts[4], pts[5]), (g, g), (pts[0], pts[0]), (pts[1], neg(pts[1])),
           (IDENTITY, pts[2]), (pts[3], IDENTITY), (IDENTITY, IDENTITY), (TWO_TORSION, pts[4]),
           (TWO_TORSION, TWO_TORSION), (pts[5], other_rep(pts[5])), (other_rep(pts[0]), pts[1])]
    return out


@pytest.mark.parametrize("code", [0, 2, 3])
def test_addition_forms(hc, code):
    rnd = random.Random(31 + code)
    for a, b in pairs(rnd):
        assert op(hc, code, a, b, rnd) == O.bw_add(a, b), (a.hex(), b.hex())


@pytest.mark.parametrize("code", [1, 4])
def test_doubling_forms(hc, code):
    rnd = random.Random(41 + code)
    for a, _ in pairs(rnd):
        assert op(hc, code, a, IDENTITY, rnd) == O.bw_dbl(a)


def test_negation(hc):
    rnd = random.Random(51)
    for a, b in pairs(rnd):
        assert op(hc, 5, a, b, rnd) == O.bw_add(neg(a), b)


def test_long_chains(hc):
    """k G by the library's doubling and mixed addition, double-and-add, for
    random 253-bit k, against the oracle's multiple."""
    rnd = random.Random(61)
    g = O.bw_generator()
    for _ in range(3):
        k = rnd.randrange(1, R_BW)
        acc = IDENTITY
        for bit in reversed(range(k.bit_length())):
            acc = op(hc, 4, acc, IDENTITY, rnd)
            if (k >> bit) & 1:
                acc = op(hc, 2, acc, g, rnd)
        assert acc == O.bw_mul_gen(k)


def test_euler_criterion(hc):
    rnd = random.Random(71)
    for _ in range(40):
        a = rnd.randrange(Q)
        want = a != 0 and pow(a, (Q - 1) // 2, Q) == 1
        assert hc.hc_bw_is_square(a.to_bytes(32, "big")) == int(want)
    assert hc.hc_bw_is_square(bytes(32)) == 0
    assert hc.hc_bw_is_square((1).to_bytes(32, "big")) == 1
    assert hc.hc_bw_is_square((7).to_bytes(32, "big")) == 0  # 7 generates the multiplicative group


def test_lazy_mixed_addition_chains(hc):
    """The accumulation loop's mixed additions (ed_madd_lazy): chains of 1, 8
    and 40 additions from the largest accepted representative of the start
    point, no canonical step in between, every coordinate checked below 2q
    after each addition; the chains include the identity, (0, -1), a point
    and its negation, the point itself (doubling) and (-x, -y)."""
    rnd = random.Random(81)
    start = O.bw_mul_gen(rnd.randrange(1, R_BW))
    for n in (1, 8, 40):
        pts = [O.bw_mul_gen(rnd.randrange(1, R_BW)) for _ in range(n)]
        if n >= 8:
            pts[1] = IDENTITY
            pts[2] = TWO_TORSION
            pts[3] = neg(pts[0])
            pts[5] = other_rep(pts[4])
        for s0 in (start, IDENTITY, pts[0]):
            za = rnd.randrange(1, Q).to_bytes(32, "little")
            out = ctypes.create_string_buffer(64)
            rc = hc.hc_bw_madd_chain(s0, za, b"".join(pts), n, out)
            assert rc == 1, rc
            want = s0
            for p in pts:
                want = O.bw_add(want, p)
            assert out.raw == want


@pytest.mark.parametrize("inline_mul", [0, 1])
def test_lazy_quad_chains(hc, inline_mul):
    """The lazily reduced quad doubling and addition (bucket reduction and
    window combine): chains of 2P then P + Q over four host lanes from the
    largest accepted representatives, no canonical step, every coordinate
    checked below 2q; with Q the identity, (0, -1), -P-like and random points."""
    rnd = random.Random(91 + inline_mul)
    g = O.bw_generator()
    cases = [(O.bw_mul_gen(rnd.randrange(1, R_BW)), O.bw_mul_gen(rnd.randrange(1, R_BW)), 24),
             (g, IDENTITY, 8), (g, TWO_TORSION, 8), (IDENTITY, g, 6), (other_rep(g), neg(g), 5),
             (TWO_TORSION, TWO_TORSION, 4)]
    for a, b, rounds in cases:
        za = rnd.randrange(1, Q).to_bytes(32, "little")
        zb = rnd.randrange(1, Q).to_bytes(32, "little")
        out = ctypes.create_string_buffer(64)
        rc = hc.hc_bw_quad_lazy_chain(a, b, za, zb, rounds, inline_mul, out)
        assert rc == 1, rc
        want = a
        for _ in range(rounds):
            want = O.bw_add(O.bw_dbl(want), b)
        assert out.raw == want


def test_row_chains_match_quad(hc):
    """The row-distributed Edwards doubling and addition (field_row.h, the
    window combine's form) in the chains of test_lazy_quad_chains: after every
    step the coordinates equal the lazy quad forms' limb for limb, in every
    row; the end point against the oracle."""
    rnd = random.Random(93)
    g = O.bw_generator()
    cases = [(O.bw_mul_gen(rnd.randrange(1, R_BW)), O.bw_mul_gen(rnd.randrange(1, R_BW)), 24),
             (g, IDENTITY, 8), (g, TWO_TORSION, 8), (IDENTITY, g, 6), (other_rep(g), neg(g), 5),
             (TWO_TORSION, TWO_TORSION, 4)]
    for a, b, rounds in cases:
        za = rnd.randrange(1, Q).to_bytes(32, "little")
        zb = rnd.randrange(1, Q).to_bytes(32, "little")
        out = ctypes.create_string_buffer(64)
        assert hc.hc_bw_row_chain(a, b, za, zb, rounds, out) == 1
        want = a
        for _ in range(rounds):
            want = O.bw_add(O.bw_dbl(want), b)
        assert out.raw == want
