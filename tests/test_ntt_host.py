"""The NTT's lazy butterflies (crypto_amd/csrc/ntt_math.h) on the CPU.

The device pass keeps values unreduced between stages (sums a + b,
differences a + 2r - b, Montgomery products without the final subtraction)
and reduces only at the end of a pass.  The host check runs the same
functions in the kernel's stage order over a whole pass of up to 2^11 points
(tests/hostcheck/hostcheck.cpp: ntt_lazy_pass), fed the widest inputs a pass
can receive (values up to 2r - 1, as a previous pass leaves them), and the
outputs must equal the oracle's transform of the inputs mod r.  The largest
intermediate must respect the stated bound (2 log2 L + 2) r, and the
quotient reduction closing a forward pass is checked on its own over the
whole range it accepts.
"""
import ctypes
import os
import random
import subprocess

import pytest

import oracle_lib as O
import pymath as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "crypto_amd", "lib", "libkinet_hostcheck.so")
FR = {0: M.BLS_R, 1: M.BN_R}


@pytest.fixture(scope="module")
def hc():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "hostcheck")], check=True)
    return ctypes.CDLL(LIB)


def le(v: int, n: int = 32) -> bytes:
    return v.to_bytes(n, "little")


def lazy_pass(hc, field, vals, inverse):
    r = FR[field]
    n = len(vals)
    lg = n.bit_length() - 1
    w = O.root_of_unity(field, n)
    if inverse:
        w = pow(w, r - 2, r)
    tw = b"".join(le(pow(w, e, r)) for e in range(max(1, n // 2)))
    out = ctypes.create_string_buffer(32 * n)
    top = ctypes.c_uint32(0)
    hc.hc_ntt_lazy_pass(field, lg, int(inverse), b"".join(le(v) for v in vals), tw, le(pow(n, r - 2, r)), out,
                        ctypes.byref(top))
    return out.raw, top.value


@pytest.mark.parametrize("field", [0, 1])
@pytest.mark.parametrize("lg", list(range(0, 12)))
def test_lazy_pass_matches_oracle(hc, field, lg):
    r = FR[field]
    n = 1 << lg
    rnd = random.Random(100 * field + lg)
    for kind in ("canonical", "up_to_2r", "max"):
        if kind == "canonical":
            vals = [rnd.randrange(r) for _ in range(n)]
        elif kind == "up_to_2r":
            vals = [rnd.randrange(2 * r) for _ in range(n)]
        else:
            vals = [2 * r - 1 - rnd.randrange(4) for _ in range(n)]
        ref_in = b"".join(le(v % r) for v in vals)
        for inverse in (False, True):
            got, top = lazy_pass(hc, field, vals, inverse)
            assert got == O.ntt(field, ref_in, inverse=inverse), (kind, inverse)
            assert top <= (((2 * lg + 2) * r - 1) >> 232), (kind, inverse, top)


@pytest.mark.parametrize("field", [0, 1])
def test_reduce_wide_whole_range(hc, field):
    """x mod r for x = q r + e over every quotient q < 64 with remainders
    0, 1, r - 1 and random ones, and around each multiple of r."""
    r = FR[field]
    rnd = random.Random(field)
    out = ctypes.create_string_buffer(32)
    cases = []
    for q in range(64):
        for e in (0, 1, r - 1, rnd.randrange(r), rnd.randrange(r)):
            cases.append(q * r + e)
    cases += [(1 << 261) - 1 - k for k in range(3) if (1 << 261) - 1 - k < 64 * r]
    for x in cases:
        assert x < 1 << 261
        hc.hc_reduce_wide(field, le(x, 36), out)
        assert int.from_bytes(out.raw, "little") == x % r, (x // r, x % r)
