"""GPU Montgomery batch inversion (gpukit_batch_inv_<field>_fp_hip and the
device form kinet_hip_batch_inv_fp) against Fermat inversion a^(p-2) mod p
in Python -- the contract of the reference's header (batch_inversion.h:1-15);
its CPU form does not return inverses (tests/test_oracle.py::
test_reference_batch_inversion_does_not_invert) -- over the reference
harness's inputs (gpukit/test/batch_inversion_test.cpp:23-33: random bytes,
top byte masked to 0x0F), unreduced inputs above p, sizes 1 .. 2^20, and the
contract's edge case: a zero element, or a nonzero multiple of p, gives
GPUKIT_ERR_BAD_SIZE with the output untouched, as the reference refuses a
zero element."""
import ctypes
import random

import pytest

import pymath as M

pytestmark = pytest.mark.gpu

P = {0: (1 << 256) - (1 << 32) - 977, 1: M.BN_P, 2: M.BLS_P}
NB = {0: 32, 1: 32, 2: 48}


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.lib.require_device()
    return crypto_amd


def harness_elems(rnd, curve, n):
    """The reference harness's random_le: random bytes, nonzero, top byte & 0x0F."""
    out = bytearray()
    for _ in range(n):
        b = bytearray(rnd.getrandbits(8) for _ in range(NB[curve]))
        if not (b[0] | b[1] | b[2] | b[3]):
            b[0] = 1
        b[-1] &= 0x0F
        out += b
    return bytes(out)


def check_inverses(curve, data, out):
    nb, p = NB[curve], P[curve]
    for i in range(len(data) // nb):
        a = int.from_bytes(data[nb * i:nb * i + nb], "little")
        b = int.from_bytes(out[nb * i:nb * i + nb], "little")
        assert b < p and a * b % p == 1, i


@pytest.mark.parametrize("curve", [0, 1, 2])
@pytest.mark.parametrize("n", [1, 2, 17, 64, 4096])
def test_matches_fermat(ca, curve, n):
    rnd = random.Random(1000 * curve + n)
    nb, p = NB[curve], P[curve]
    print(f"batch inversion curve={curve} n={n}", flush=True)
    for kind in ("harness", "unreduced"):
        if kind == "harness":
            data = harness_elems(rnd, curve, n)
        else:  # full-width random bytes: values up to 2^(8 nb) > p
            data = bytes(rnd.getrandbits(8) for _ in range(nb * n))
        got = ca.batch_inverse(curve, data)
        want = b"".join(pow(int.from_bytes(data[nb * i:nb * i + nb], "little") % p, p - 2, p).to_bytes(nb, "little")
                        for i in range(n))
        assert got == want, kind
        # in place (the host buffers may be one)
        buf = ctypes.create_string_buffer(data, len(data))
        fn = getattr(ca.library(), ca.lib._BATCH_INV[curve])
        assert fn(buf, buf, n) == 0 and buf.raw[: len(data)] == want


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_full_size_inverses(ca, curve):
    """a * inv == 1 mod p: every element at 2^17 (the reference harness's
    largest size, its inputs), a sample of 4096 at 2^20 (full-width
    unreduced inputs)."""
    rnd = random.Random(77 + curve)
    for lg in (17, 20):
        n = 1 << lg
        print(f"batch inversion curve={curve} n=2^{lg}", flush=True)
        data = harness_elems(rnd, curve, n) if lg == 17 else rnd.randbytes(NB[curve] * n)
        got = ca.batch_inverse(curve, data)
        nb, p = NB[curve], P[curve]
        idx = range(n) if lg == 17 else rnd.sample(range(n), 4096)
        for i in idx:
            a = int.from_bytes(data[nb * i:nb * i + nb], "little")
            b = int.from_bytes(got[nb * i:nb * i + nb], "little")
            assert b < p and a * b % p == 1, (lg, i)


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_zero_elements_are_refused(ca, curve):
    """A zero element anywhere, or a nonzero multiple of p (p, 2p where it
    fits the record): GPUKIT_ERR_BAD_SIZE and the output untouched."""
    import crypto_amd

    lib = crypto_amd.library()
    nb, p = NB[curve], P[curve]
    rnd = random.Random(5 + curve)
    base = bytearray(harness_elems(rnd, curve, 300))
    fn = getattr(lib, crypto_amd.lib._BATCH_INV[curve])
    zeros = [bytes(nb)] + [(m * p).to_bytes(nb, "little") for m in (1, 2) if m * p < 1 << (8 * nb)]
    for z in zeros:
        for pos in (0, 123, 299):
            data = bytearray(base)
            data[nb * pos:nb * pos + nb] = z
            out = ctypes.create_string_buffer(b"\x5a" * len(data), len(data))
            assert fn(bytes(data), out, 300) == -2
            assert out.raw == b"\x5a" * len(data)
    assert fn(None, None, 0) == 0  # n = 0: OK, nothing read or written


@pytest.mark.parametrize("curve", [1, 2])
def test_device_form_status_and_graph(ca, curve):
    """kinet_hip_batch_inv_fp on device buffers: the host form's results, the
    status word, the overlap refusal, and the same results from a captured
    graph replayed with eager work between replays."""
    import torch

    import crypto_amd

    lib = crypto_amd.library()
    nb = NB[curve]
    rnd = random.Random(9 + curve)
    n = 5000
    data = harness_elems(rnd, curve, n)
    want = ca.batch_inverse(curve, data)
    d_in = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
    d_out = torch.empty_like(d_in)
    d_st = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    assert lib.kinet_hip_batch_inv_fp(curve, d_in.data_ptr(), d_out.data_ptr(), n, d_st.data_ptr(),
                                      ctypes.c_void_p(s.cuda_stream)) == 0
    s.synchronize()
    assert int(d_st.item()) == 0 and bytes(d_out.cpu().numpy()) == want
    assert lib.kinet_hip_batch_inv_fp(curve, d_in.data_ptr(), d_in.data_ptr() + nb, n, d_st.data_ptr(),
                                      ctypes.c_void_p(s.cuda_stream)) == -2
    # status: 1 for a zero element or a multiple of p
    for z in (bytes(nb), P[curve].to_bytes(nb, "little")):
        bad = bytearray(data)
        bad[nb * 17:nb * 18] = z
        d_bad = torch.frombuffer(bad, dtype=torch.uint8).cuda()
        assert lib.kinet_hip_batch_inv_fp(curve, d_bad.data_ptr(), d_out.data_ptr(), n, d_st.data_ptr(),
                                          ctypes.c_void_p(s.cuda_stream)) == 0
        s.synchronize()
        assert int(d_st.item()) == 1
    # captured once, replayed three times with eager work in between
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        assert lib.kinet_hip_batch_inv_fp(curve, d_in.data_ptr(), d_out.data_ptr(), n, d_st.data_ptr(),
                                          ctypes.c_void_p(s.cuda_stream)) == 0
    for rep in range(3):
        d_out.fill_(0x33)
        d_st.fill_(9)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(d_st.item()) == 0 and bytes(d_out.cpu().numpy()) == want, rep
