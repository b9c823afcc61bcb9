"""The library's field and curve arithmetic (crypto_amd/csrc/field.h, curve.h),
compiled for the host, against Python integers and the oracle.  Catches
limb/radix/reduction slips before any GPU run."""
import ctypes
import os
import random
import subprocess

import pytest

import oracle_lib as O
import pymath as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "crypto_amd", "lib", "libkinet_hostcheck.so")


@pytest.fixture(scope="module")
def hc():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "hostcheck")], check=True)
    return ctypes.CDLL(LIB)


FIELDS = {0: (M.BLS_P, 48), 1: (M.BLS_R, 32), 2: (M.BN_P, 32), 3: (M.BN_R, 32), 4: (M.SECP_P, 32)}


@pytest.mark.parametrize("field", [0, 1, 2, 3, 4])
def test_field_ops(hc, field):
    p, nb = FIELDS[field]
    rnd = random.Random(100 + field)
    special = [0, 1, 2, p - 1, p - 2, (p - 1) // 2, (p + 1) // 2, (1 << 29) - 1, 1 << 29, p >> 29 << 29]
    vals = special + [rnd.randrange(p) for _ in range(200)]
    out = ctypes.create_string_buffer(nb)
    for i, a in enumerate(vals):
        b = vals[(i * 7 + 3) % len(vals)] if i < len(special) else rnd.randrange(p)
        A, B = a.to_bytes(nb, "little"), b.to_bytes(nb, "little")
        for op, want in ((0, (a + b) % p), (1, (a - b) % p), (2, a * b % p), (4, a * a % p)):
            hc.hc_field_op(field, op, A, B, out)
            assert int.from_bytes(out.raw, "little") == want, (field, op, a, b)
        if a:
            hc.hc_field_op(field, 3, A, B, out)
            assert int.from_bytes(out.raw, "little") == pow(a, p - 2, p), (field, a)


@pytest.mark.parametrize("field", [0, 1, 2, 3, 4])
def test_field_inverse_many(hc, field):
    """Divstep inversion against Fermat's a^(p-2) over random elements and
    inputs whose gcd runs are long or short: powers of two, 2^k - 1, values
    next to p and p/2, small integers."""
    p, nb = FIELDS[field]
    rnd = random.Random(700 + field)
    vals = [1 << k for k in range(p.bit_length())] + [(1 << k) - 1 for k in range(2, p.bit_length())]
    vals += [p - k for k in range(1, 40)] + [p // 2 + k for k in range(-5, 6)] + list(range(1, 60))
    vals += [rnd.randrange(1, p) for _ in range(1500)]
    out = ctypes.create_string_buffer(nb)
    zero = bytes(nb)
    for a in vals:
        a %= p
        if not a:
            continue
        hc.hc_field_op(field, 3, a.to_bytes(nb, "little"), zero, out)
        assert int.from_bytes(out.raw, "little") == pow(a, p - 2, p), (field, a)


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_lazy_mixed_addition_chain(hc, curve):
    """Mixed additions of the accumulation loop (coordinates kept below 8p /
    4p / 2p between them), then made canonical: P + 3Q for random points, for
    Q = P (the doubling case on the first step), Q = -P (the sum passes
    through the identity), P = -2Q and P = -3Q (identity mid-chain and at the
    end) and P the identity; and P + 20Q, where the bounds reach their steady
    state, including chains through the identity."""
    pb = O.POINT_BYTES[curve]
    rnd = random.Random(40 + curve)
    out = ctypes.create_string_buffer(pb)
    G = M.CURVES[curve]["g"]
    r = M.CURVES[curve]["r"]
    cases = [(rnd.randrange(1, r), rnd.randrange(1, r)) for _ in range(12)]
    q = rnd.randrange(1, r)
    cases += [(q, q), (r - q, q), ((-2 * q) % r, q), ((-3 * q) % r, q), (0, q), ((-1 * q) % r, q)]
    for kp, kq in cases:
        P = M.mul(curve, G, kp) if kp else None
        Q = M.mul(curve, G, kq)
        Pw = M.to_wire(curve, P) if P is not None else bytes(pb)
        hc.hc_point_op(curve, 3, Pw, M.to_wire(curve, Q), out)
        want = (kp + 3 * kq) % r
        assert out.raw == (M.to_wire(curve, M.mul(curve, G, want)) if want else bytes(pb)), (kp, kq)
    for kp, kq in cases[:6] + [((-20 * q) % r, q), ((-7 * q) % r, q)]:  # 20 additions; identity mid-chain
        P = M.mul(curve, G, kp) if kp else None
        Pw = M.to_wire(curve, P) if P is not None else bytes(pb)
        hc.hc_point_op(curve, 5, Pw, M.to_wire(curve, M.mul(curve, G, kq)), out)
        want = (kp + 20 * kq) % r
        assert out.raw == (M.to_wire(curve, M.mul(curve, G, want)) if want else bytes(pb)), (kp, kq, 20)


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_point_ops(hc, curve):
    pb = O.POINT_BYTES[curve]
    rnd = random.Random(curve)
    out = ctypes.create_string_buffer(pb)
    G = M.CURVES[curve]["g"]
    inf = bytes(pb)
    for t in range(12):
        P = M.mul(curve, G, rnd.getrandbits(200))
        Q = M.mul(curve, G, rnd.getrandbits(200)) if t % 3 else P
        Pw, Qw = M.to_wire(curve, P), M.to_wire(curve, Q)
        for op in (0, 1):
            hc.hc_point_op(curve, op, Pw, Qw, out)
            assert out.raw == M.to_wire(curve, M.add(curve, P, Q)), (t, op)
        hc.hc_point_op(curve, 2, Pw, Pw, out)
        assert out.raw == M.to_wire(curve, M.add(curve, P, P))
        negQ = M.to_wire(curve, M.neg(curve, Q))
        hc.hc_point_op(curve, 0, Qw, negQ, out)
        assert out.raw == inf
        hc.hc_point_op(curve, 1, Qw, negQ, out)
        assert out.raw == inf
        hc.hc_point_op(curve, 0, inf, Qw, out)
        assert out.raw == Qw
        hc.hc_point_op(curve, 1, inf, Qw, out)
        assert out.raw == Qw


BN_LAMBDA = 0x30644E72E131A029048B6E193FD84104CC37A73FEC2BC5E9B8CA0B2D36636F23


def test_glv_split_bn(hc):
    """The BN254 GLV split of the library (host build): s = k1 + lambda k2 mod r
    with |k1|, |k2| < 2^127 for random and edge scalars, lambda a cube root of
    unity mod r; and lambda G = phi(G) = (beta x_G, y_G) on the curve."""
    r = M.CURVES[1]["r"]
    p = M.CURVES[1]["p"]
    assert pow(BN_LAMBDA, 3, r) == 1 and BN_LAMBDA != 1
    beta = 0x30644E72E131A0295E6DD9E7E0ACCCB0C28F069FBB966E3DE4BD44E5607CFD48
    G = M.CURVES[1]["g"]
    assert M.mul(1, G, BN_LAMBDA) == (beta * G[0] % p, G[1])
    rnd = random.Random(5)
    out = ctypes.create_string_buffer(34)
    edge = [0, 1, 2, r - 1, r, r + 1, (1 << 256) - 1, BN_LAMBDA, r - BN_LAMBDA, (r - 1) // 2, 4 * r, 5 * r - 1]
    for s in edge + [rnd.getrandbits(256) for _ in range(3000)]:
        hc.hc_glv_split_bn(s.to_bytes(32, "little"), out)
        raw = out.raw
        k1 = int.from_bytes(raw[:16], "little") * (-1 if raw[32] else 1)
        k2 = int.from_bytes(raw[16:32], "little") * (-1 if raw[33] else 1)
        assert abs(k1) < 1 << 127 and abs(k2) < 1 << 127, hex(s)
        assert (k1 + BN_LAMBDA * k2 - s) % r == 0, hex(s)



SECP_LAMBDA = 0x5363AD4CC05C30E0A5261C028812645A122E22EA20816678DF02967C1B23BD72
SECP_BETA = 0x7AE96A2B657C07106E64479EAC3434E99CF0497512F58995C1396C28719501EE


def test_glv_split_secp(hc):
    """The secp256k1 GLV split (host build): s = k1 + lambda k2 mod n with
    |k1| <= (A1 + A2)/2 + 1 and |k2| <= (B1 + B2)/2 + 1, both below 2^128, for
    random and edge scalars (0, n - 1, n, 2^256 - 1, lambda and the basis
    entries); lambda is a cube root of unity mod n with lambda G = (beta x_G, y_G),
    and the basis vectors lie in the lattice {(a, b): a + b lambda = 0 mod n}
    with determinant n."""
    n = M.CURVES[0]["r"]
    p = M.CURVES[0]["p"]
    lam = SECP_LAMBDA
    assert pow(lam, 3, n) == 1 and lam != 1 and pow(SECP_BETA, 3, p) == 1
    G = M.CURVES[0]["g"]
    assert M.mul(0, G, lam) == (SECP_BETA * G[0] % p, G[1])
    a1, b1 = 0x3086D221A7D46BCDE86C90E49284EB15, 0xE4437ED6010E88286F547FA90ABFE4C3
    a2, b2 = 0x114CA50F7A8E2F3F657C1108D9D44CFD8, 0x3086D221A7D46BCDE86C90E49284EB15
    assert (a1 - b1 * lam) % n == 0 and (a2 + b2 * lam) % n == 0 and a1 * b2 + b1 * a2 == n
    lim1, lim2 = (a1 + a2) // 2 + 1, (b1 + b2) // 2 + 1
    rnd = random.Random(7)
    out = ctypes.create_string_buffer(34)
    edge = [0, 1, 2, n - 1, n, n + 1, (1 << 256) - 1, lam, n - lam, a1, a2, b1, b2, (n - 1) // 2, n // 3]
    max_bits = 0
    for s in edge + [rnd.getrandbits(256) for _ in range(4000)] + [rnd.randrange(n) for _ in range(2000)]:
        hc.hc_glv_split_secp(s.to_bytes(32, "little"), out)
        raw = out.raw
        k1 = int.from_bytes(raw[:16], "little") * (-1 if raw[32] else 1)
        k2 = int.from_bytes(raw[16:32], "little") * (-1 if raw[33] else 1)
        assert abs(k1) <= lim1 and abs(k2) <= lim2, hex(s)
        assert (k1 + lam * k2 - s) % n == 0, hex(s)
        max_bits = max(max_bits, abs(k1).bit_length(), abs(k2).bit_length())
    assert max_bits <= 128


def test_glv_split_bls(hc):
    """The BLS12-381 signed GLV split (host build): s = k1 + lambda k2 mod r with
    |k1|, |k2| < 2^127 (the magnitude bound that keeps 16-bit windows free of
    carries into a ninth window), for random scalars, for the boundaries of
    both moves (k2 near (lambda + 1)/2, k1 near lambda/2, s = lambda^2 +
    lambda) and for scalars of 2^256 - 1 and multiples of r; r = lambda^2 +
    lambda + 1 and lambda G = phi(G) pin the constants."""
    lam = 0xAC45A4010001A40200000000FFFFFFFF
    r = M.CURVES[2]["r"]
    p = M.CURVES[2]["p"]
    assert r == lam * lam + lam + 1
    G = M.CURVES[2]["g"]
    lg = M.mul(2, G, lam)
    assert lg[1] == G[1] and pow(lg[0] * pow(G[0], -1, p) % p, 3, p) == 1
    rnd = random.Random(6)
    out = ctypes.create_string_buffer(34)
    h1, h = (lam + 1) // 2, lam // 2
    edge = [0, 1, 2, r - 1, r, r + 1, (1 << 256) - 1, lam, lam - 1, lam + 1, lam * lam + lam, lam * lam + lam - 1,
            h, h + 1, h - 1, h1 * lam, (h1 + 1) * lam, (h1 + 1) * lam - 1, (h1 + 1) * lam + h, (h1 + 1) * lam + h + 1,
            h1 * lam + h + 1, 2 * r, 3 * r - 1, 2 * r + lam]
    edge += [q * lam + k for q in (0, 1, h1 - 1, h1, h1 + 1, lam - 1, lam) for k in (0, 1, h - 1, h, h + 1, lam - 1)]
    big = max_bits = 0
    for s in edge + [rnd.getrandbits(256) for _ in range(4000)] + [rnd.randrange(r) for _ in range(2000)]:
        s %= 1 << 256
        hc.hc_glv_split_bls(s.to_bytes(32, "little"), out)
        raw = out.raw
        k1 = int.from_bytes(raw[:16], "little") * (-1 if raw[32] else 1)
        k2 = int.from_bytes(raw[16:32], "little") * (-1 if raw[33] else 1)
        assert abs(k1) <= h and abs(k2) <= h1 + 1, hex(s)
        assert (k1 + lam * k2 - s) % r == 0, hex(s)
        max_bits = max(max_bits, abs(k1).bit_length(), abs(k2).bit_length())
    assert max_bits <= 127

# ------------------------------------------------ quad-cooperative operations
# xyzz_add_quad / xyzz_dbl_quad (curve_quad.h) are the fold, reduction and
# combine's point operations.  Here they run on the host, unchanged, with four
# lock-stepped threads as the quad's lanes.  Inputs are points in the lazy
# forms the device passes between operations: ZZ = z^2, ZZZ = z^3 for a
# random z, and each coordinate raised by a multiple of p up to the lazy
# bounds (X < 8p, Y < 4p, ZZ, ZZZ < 2p).  BN254's headroom R / p is about 169
# against 2^25 for BLS12-381, so bound slips show on BN254 first.
LAZY_MAX = (7, 3, 1, 1)


def _rand_points(curve, count, rnd):
    """count random points of the group as (x, y): k G by the C oracle (BN254,
    BLS12-381), or a progression P + i D for secp256k1 (which it lacks)."""
    G, r = M.CURVES[curve]["g"], M.CURVES[curve]["r"]
    if curve == 0:
        P, D = M.mul(0, G, rnd.randrange(1, r)), M.mul(0, G, rnd.randrange(1, r))
        out = []
        for _ in range(count):
            out.append(P)
            P = M.add(0, P, D)
        return out
    return [M.from_wire(curve, O.g1_mul_gen(curve, rnd.randrange(1, r).to_bytes(32, "little")))
            for _ in range(count)]


def _quad_args(curve, P, rnd, lazy, extreme=False):
    p, fb = M.CURVES[curve]["p"], M.CURVES[curve]["fb"]
    z = rnd.randrange(1, p)
    if extreme:
        k = LAZY_MAX if lazy else (0, 0, 0, 0)
    else:
        k = tuple(rnd.randint(0, m) for m in LAZY_MAX) if lazy else (0, 0, 0, 0)
    w = M.to_wire(curve, P) if P is not None else bytes(2 * fb)
    return w, z.to_bytes(fb, "little"), (ctypes.c_int * 4)(*k)


@pytest.mark.parametrize("curve", [0, 1, 2])
@pytest.mark.parametrize("lazy", [0, 1])
def test_quad_ops_match_textbook(hc, curve, lazy):
    """a + b and 2a by the quad operations for random points, a = b (the
    doubling case reached through the addition, with different z, so the
    inputs differ as XYZZ and agree only as points), a = -b, identities on
    either side; coordinates at random and at the extreme lazy offsets."""
    G, r = M.CURVES[curve]["g"], M.CURVES[curve]["r"]
    pb = O.POINT_BYTES[curve]
    rnd = random.Random(1000 + 10 * curve + lazy)
    out = ctypes.create_string_buffer(pb)
    pts = _rand_points(curve, 24, rnd)
    cases = [(pts[i], pts[i + 1]) for i in range(0, 16, 2)]
    cases += [(pts[16], pts[16]), (pts[17], M.neg(curve, pts[17])), (None, pts[18]), (pts[19], None), (None, None)]
    cases += [(pts[20], M.add(curve, pts[20], pts[20]))]
    for i, (A, B) in enumerate(cases):
        for extreme in (False, True):
            wa, za, ka = _quad_args(curve, A, rnd, lazy, extreme)
            wb, zb, kb = _quad_args(curve, B, rnd, lazy, extreme)
            assert hc.hc_quad_op(curve, 0, lazy, wa, wb, za, zb, ka, kb, out) == 0, "lanes disagree"
            want = M.add(curve, A, B)
            assert out.raw == (M.to_wire(curve, want) if want is not None else bytes(pb)), (curve, lazy, i, extreme)
            assert hc.hc_quad_op(curve, 1, lazy, wa, wa, za, za, ka, ka, out) == 0
            want = M.add(curve, A, A)
            assert out.raw == (M.to_wire(curve, want) if want is not None else bytes(pb)), (curve, lazy, i, "dbl")


@pytest.mark.parametrize("curve", [0, 1, 2])
@pytest.mark.parametrize("n", [8, 64])
def test_quad_lazy_trees(hc, curve, n):
    """Trees of lazy quad additions with no canonical step between levels
    (fan-in 8 has three levels, fan-in 64 six) over random points with
    identities, repeated points and cancelling pairs mixed in, so doublings
    and identities occur mid-tree on lazy inputs; against the textbook sum."""
    G, r, p, fb = M.CURVES[curve]["g"], M.CURVES[curve]["r"], M.CURVES[curve]["p"], M.CURVES[curve]["fb"]
    pb = O.POINT_BYTES[curve]
    rnd = random.Random(2000 + curve * 100 + n)
    out = ctypes.create_string_buffer(pb)
    for trial in range(6):
        pts = _rand_points(curve, n, rnd)
        if trial >= 1:
            pts[1] = None
            pts[n // 2] = pts[0]                      # doubling at the first level
            pts[n // 2 + 1] = M.neg(curve, pts[1] or pts[2])
        if trial >= 3:
            pts = [pts[0]] * n                         # every level doubles
        if trial == 5:
            pts = [pts[0], M.neg(curve, pts[0])] * (n // 2)  # sums through the identity
        wire = b"".join(M.to_wire(curve, P) if P is not None else bytes(2 * fb) for P in pts)
        zs = b"".join(rnd.randrange(1, p).to_bytes(fb, "little") for _ in range(n))
        want = None
        for P in pts:
            want = M.add(curve, want, P)
        for lazy in (1, 0):
            assert hc.hc_quad_tree(curve, lazy, wire, zs, n, out) == 0, "lanes disagree"
            assert out.raw == (M.to_wire(curve, want) if want is not None else bytes(pb)), (curve, n, trial, lazy)


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_quad_doubling_chains(hc, curve):
    """Chains of lazy quad doublings as in the window combine (16 per window,
    112 over a 127-bit GLV scalar), from points at the extreme lazy offsets."""
    G, r = M.CURVES[curve]["g"], M.CURVES[curve]["r"]
    pb = O.POINT_BYTES[curve]
    rnd = random.Random(3000 + curve)
    out = ctypes.create_string_buffer(pb)
    for count in (1, 16, 112):
        P = _rand_points(curve, 1, rnd)[0]
        w, z, k = _quad_args(curve, P, rnd, 1, extreme=True)
        assert hc.hc_quad_dbl_chain(curve, w, z, k, count, out) == 0
        assert out.raw == M.to_wire(curve, M.mul(curve, P, 1 << count)), (curve, count)


# ---------------------------------------------------------------- rows
ROW_N = {0: 10, 1: 9, 2: 14}


def _limbs(v, n):
    return [(v >> (29 * i)) & ((1 << 29) - 1) for i in range(n)]


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_row_mul_matches_scan(hc, curve):
    """The row-distributed Montgomery product (field_row.h, four products per
    wave, limb j in lane j of a row) against the one-lane product scan
    fe_mul_k<F, 1, false>, limb for limb, and against a b 2^(-29 N) mod p,
    below 2p: operands 0, 1, p - 1, p, 2p - 1, 8p - 1 (the lazy bound) and
    random ones below 8p."""
    p, n = M.CURVES[curve]["p"], ROW_N[curve]
    rinv = pow(1 << (29 * n), -1, p)
    rnd = random.Random(4000 + curve)
    special = [0, 1, p - 1, p, 2 * p - 1, 8 * p - 1, (1 << 29) - 1, p >> 29 << 29]
    vals = special + [rnd.randrange(8 * p) for _ in range(200)]
    A, B = (ctypes.c_uint32 * 64)(), (ctypes.c_uint32 * 64)()
    out, ref = (ctypes.c_uint32 * 64)(), (ctypes.c_uint32 * 64)()
    for t in range(0, len(vals), 2):
        pairs = [(vals[(t + q) % len(vals)], vals[(t * 5 + 3 * q + 1) % len(vals)]) for q in range(4)]
        for q, (a, b) in enumerate(pairs):
            la, lb = _limbs(a, n), _limbs(b, n)
            for j in range(16):
                A[16 * q + j] = la[j] if j < n else 0
                B[16 * q + j] = lb[j] if j < n else 0
        assert hc.hc_row_mul(curve, A, B, out, ref) == 0, "lanes N..15 not zero"
        for q, (a, b) in enumerate(pairs):
            got = list(out[16 * q:16 * q + n])
            assert got == list(ref[16 * q:16 * q + n]), (curve, a, b)
            v = sum(x << (29 * i) for i, x in enumerate(got))
            assert v % p == a * b * rinv % p and v < 2 * p, (curve, a, b)


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_row_ops_match_quad(hc, curve):
    """Row-distributed a + b and 2a (field_row.h) against the lazy quad
    operations, coordinate limb for coordinate limb, and against the textbook
    result: random points, a = b through the addition, a = -b, identities,
    coordinates at random and at the extreme lazy offsets."""
    pb = O.POINT_BYTES[curve]
    n = ROW_N[curve]
    rnd = random.Random(5000 + curve)
    out = ctypes.create_string_buffer(pb)
    raw_r, raw_q = (ctypes.c_uint32 * (4 * n))(), (ctypes.c_uint32 * (4 * n))()
    pts = _rand_points(curve, 24, rnd)
    cases = [(pts[i], pts[i + 1]) for i in range(0, 16, 2)]
    cases += [(pts[16], pts[16]), (pts[17], M.neg(curve, pts[17])), (None, pts[18]), (pts[19], None), (None, None)]
    cases += [(pts[20], M.add(curve, pts[20], pts[20]))]
    for i, (A, B) in enumerate(cases):
        for extreme in (False, True):
            wa, za, ka = _quad_args(curve, A, rnd, 1, extreme)
            wb, zb, kb = _quad_args(curve, B, rnd, 1, extreme)
            for op, want in ((0, M.add(curve, A, B)), (1, M.add(curve, A, A))):
                assert hc.hc_row_op(curve, op, wa, wb, za, zb, ka, kb, out, raw_r, raw_q) == 0, "rows disagree"
                assert list(raw_r) == list(raw_q), (curve, i, extreme, op)
                assert out.raw == (M.to_wire(curve, want) if want is not None else bytes(pb)), (curve, i, op)


@pytest.mark.parametrize("curve", [0, 1, 2])
@pytest.mark.parametrize("nwin,c", [(8, 16), (3, 5), (15, 9)])
def test_row_combine(hc, curve, nwin, c):
    """The window combine by rows, sum_w 2^(c w) S_w by Horner from the top,
    over window sums at the extreme lazy offsets; cases with identity
    windows (the top one included), a window equal to the running sum (the
    addition's doubling case) and one cancelling it (the identity mid-chain)."""
    G, r, p, fb = M.CURVES[curve]["g"], M.CURVES[curve]["r"], M.CURVES[curve]["p"], M.CURVES[curve]["fb"]
    pb = O.POINT_BYTES[curve]
    rnd = random.Random(6000 + 100 * curve + nwin)
    out = ctypes.create_string_buffer(pb)
    for trial in range(4):
        ks = [rnd.randrange(r) for _ in range(nwin)]
        if trial == 1:
            ks[nwin - 1] = 0
            ks[nwin // 2] = 0
        if trial >= 2:
            # running value before window w: sum_{v > w} 2^(c (v - w)) k_v
            w = nwin // 2
            run = sum(ks[v] << (c * (v - w)) for v in range(w + 1, nwin)) % r
            ks[w] = run if trial == 2 else (r - run) % r
        want = sum(k << (c * w) for w, k in enumerate(ks)) % r
        wins = b"".join(M.to_wire(curve, M.mul(curve, G, k)) if k else bytes(2 * fb) for k in ks)
        zs = b"".join(rnd.randrange(1, p).to_bytes(fb, "little") for _ in range(nwin))
        kk = (ctypes.c_int * (4 * nwin))(*(LAZY_MAX * nwin))
        assert hc.hc_row_combine(curve, wins, zs, kk, nwin, c, out) == 0, "rows disagree"
        assert out.raw == (M.to_wire(curve, M.mul(curve, G, want)) if want else bytes(pb)), (curve, trial)
