"""G1 decompression on the device (kinet_hip_points_decompress) against the
oracle's orc_g1_decompress: the same records for valid encodings (round trip
of compressed points, identity included) and the same rejections (missing
compression flag, loose infinity, x >= p, x^3 + b not a square)."""
import random

import pytest
import torch

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    import crypto_amd
    from crypto_amd import device

    crypto_amd.library()
    return device


def _dev(b):
    return torch.frombuffer(bytearray(b), dtype=torch.uint8).to("cuda")


@pytest.mark.parametrize("curve", [1, 2])
def test_decompress_round_trip(D, curve):
    n = 4096 + 37
    pb, cb = O.POINT_BYTES[curve], O.COMP_BYTES[curve]
    wire = bytearray(O.gen_points_progression(curve, 0x51ED, 0xFACE5, n))
    for i in range(3, n, 613):
        wire[i * pb:(i + 1) * pb] = bytes(pb)  # identity points
    comp = b"".join(O.compress(curve, bytes(wire[i * pb:(i + 1) * pb])) for i in range(n))
    print(f"decompress curve={curve} n={n}", flush=True)
    out, ok = D.points_decompress(curve, _dev(comp))
    assert ok
    assert out.cpu().numpy().tobytes() == bytes(wire)


def _invalid_encodings(curve):
    p, b = M.CURVES[curve]["p"], M.CURVES[curve]["b"]
    cb = O.COMP_BYTES[curve]
    top = 0x80
    bad = []
    G = M.CURVES[curve]["g"]
    good = bytearray(O.compress(curve, M.to_wire(curve, M.mul(curve, G, 12345))))
    nf = bytearray(good)
    nf[0] &= 0x3F if curve == 1 else 0x7F  # no compression flag
    bad.append(bytes(nf))
    flag = 0xC0 if curve == 2 else 0x40
    bad.append(bytes([flag]) + bytes(cb - 2) + b"\x01")  # loose infinity
    bad.append(bytes([flag | 0x01]) + bytes(cb - 1))
    xp = bytearray(p.to_bytes(cb, "big"))
    xp[0] |= top
    bad.append(bytes(xp))  # x = p
    for x in range(2, 200):  # x^3 + b not a square
        if M.sqrt(curve, (x * x * x + b) % p) is None:
            e = bytearray(x.to_bytes(cb, "big"))
            e[0] |= top
            bad.append(bytes(e))
            break
    return bytes(good), bad


@pytest.mark.parametrize("curve", [1, 2])
def test_decompress_rejections_match_oracle(D, curve):
    good, bad = _invalid_encodings(curve)
    pb = O.POINT_BYTES[curve]
    for enc in bad:
        assert O.decompress(curve, enc) is None, enc.hex()
        out, ok = D.points_decompress(curve, _dev(good + enc + good))
        assert not ok, enc.hex()
        rec = out.cpu().numpy().tobytes()
        want = O.decompress(curve, good)
        assert rec[:pb] == want and rec[2 * pb:] == want and rec[pb:2 * pb] == bytes(pb)
    out, ok = D.points_decompress(curve, _dev(good * 3))
    assert ok and out.cpu().numpy().tobytes() == O.decompress(curve, good) * 3
