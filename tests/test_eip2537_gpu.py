"""EIP-2537 G1 MSM (bls12_381_g1_msm, include/kinet/crypto/bls.h) against
the reference's semantics (bls/cpp/bls.cpp:194-245): results against the
closed form sum c_k k_k G for points P_k = k_k G, full 256-bit big-endian
scalars, infinity entries skipped, and every rejection the reference makes
(coordinate padding, coordinates >= p, off-curve and non-G1 points, sizes that
are not a multiple of 160)."""
import ctypes
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu

CURVE = 2
P = M.CURVES[CURVE]["p"]
R = M.CURVES[CURVE]["r"]
G = M.CURVES[CURVE]["g"]


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.library()
    return crypto_amd


def enc_fp(v):
    return v.to_bytes(64, "big")


def entry(point, c):
    x, y = point if point is not None else (0, 0)
    return enc_fp(x) + enc_fp(y) + c.to_bytes(32, "big")


def expect(total):
    Q = M.mul(CURVE, G, total % R) if total % R else None
    return (enc_fp(Q[0]), enc_fp(Q[1])) if Q is not None else (bytes(64), bytes(64))


@pytest.mark.parametrize("n", [1, 7, 300, 5000])
def test_g1_msm_closed_form(ca, n):
    rnd = random.Random(n)
    k0, d = 0x1234567 + n, 0xABCDEF1234567
    ks = [k0 + i * d for i in range(n)]
    cs = [rnd.getrandbits(256) for _ in range(n)]
    if n > 2:
        cs[0], cs[1] = (1 << 256) - 1, R  # scalars used in full (>= r)
    wire = O.gen_points_progression(CURVE, k0, d, n)  # (k0 + i d) G, 48-byte BE x || y
    data = b"".join(bytes(16) + wire[96 * i:96 * i + 48] + bytes(16) + wire[96 * i + 48:96 * i + 96]
                    + c.to_bytes(32, "big") for i, c in enumerate(cs))
    print(f"eip2537 msm n={n}", flush=True)
    ok, rx, ry = ca.bls12_381_g1_msm(data)
    assert ok
    assert (rx, ry) == expect(sum(c * k for c, k in zip(cs, ks)))


def test_g1_msm_infinity_and_empty(ca):
    Q = M.mul(CURVE, G, 5)
    assert ca.bls12_381_g1_msm(b"") == (True, bytes(64), bytes(64))
    assert ca.bls12_381_g1_msm(entry(None, 7) * 3) == (True, bytes(64), bytes(64))
    # infinity entries are skipped; Q and -Q cancel to infinity
    negQ = (Q[0], (P - Q[1]) % P)
    assert ca.bls12_381_g1_msm(entry(Q, 3) + entry(None, 9) + entry(negQ, 3)) == (True, bytes(64), bytes(64))
    ok, rx, ry = ca.bls12_381_g1_msm(entry(None, 1) + entry(Q, 2))
    assert ok and (rx, ry) == expect(10)


def _non_g1_point():
    b = M.CURVES[CURVE]["b"]
    for x in range(1, 1000):
        y = M.sqrt(CURVE, (x * x * x + b) % P)
        if y is not None and M.mul(CURVE, (x, y), R) is not None:
            return (x, y)
    raise AssertionError("no point found")


def test_g1_msm_rejections(ca):
    Q = M.mul(CURVE, G, 11)
    good = entry(Q, 3)
    assert ca.bls12_381_g1_msm(good)[0]
    # padding: a non-zero byte in the first 16 bytes of x or y
    for off in (0, 15, 64, 79):
        bad = bytearray(good)
        bad[off] = 1
        assert ca.bls12_381_g1_msm(bytes(bad))[0] is False, off
    # coordinates >= p (x = p with y kept, and y + p)
    assert ca.bls12_381_g1_msm(enc_fp(P) + enc_fp(Q[1]) + (3).to_bytes(32, "big"))[0] is False
    assert ca.bls12_381_g1_msm(enc_fp(Q[0]) + enc_fp(Q[1] + P) + (3).to_bytes(32, "big"))[0] is False
    # off the curve, and on the curve but outside G1 (also among valid entries)
    off_curve = (Q[0], (Q[1] + 1) % P)
    assert ca.bls12_381_g1_msm(good + entry(off_curve, 1))[0] is False
    assert ca.bls12_381_g1_msm(good * 5 + entry(_non_g1_point(), 0))[0] is False
    # sizes that are not whole entries
    assert ca.bls12_381_g1_msm(good[:-1])[0] is False
    assert ca.bls12_381_g1_msm(good + good[:10])[0] is False
    # outputs untouched on failure
    lib = ca.library()
    rx = ctypes.create_string_buffer(b"\x55" * 64, 64)
    ry = ctypes.create_string_buffer(b"\x55" * 64, 64)
    assert lib.bls12_381_g1_msm(rx, ry, good[:-1], len(good) - 1) == 0
    assert rx.raw == b"\x55" * 64 and ry.raw == b"\x55" * 64
