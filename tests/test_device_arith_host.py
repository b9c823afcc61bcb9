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


FIELDS = {0: (M.BLS_P, 48), 1: (M.BLS_R, 32), 2: (M.BN_P, 32), 3: (M.BN_R, 32)}


@pytest.mark.parametrize("field", [0, 1, 2, 3])
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


@pytest.mark.parametrize("curve", [1, 2])
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
