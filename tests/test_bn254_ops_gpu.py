"""kinet_bn254_hip_{g1_add, g1_mul, fp_mul} (include/kinet/crypto/bn254.h)
through the C ABI, in the reference driver's batch layout:
  - byte for byte against the reference's own CPU path of
    kinet_bn254_cuda_{g1_add, g1_mul, fp_mul} (oracle/_ref/libbnref.so) on 100
    deterministic vectors per operation, the pattern of the reference's
    determinism harness (bn254/test/bn254_gpu_determinism_test.cpp);
  - against the golden products and point operations the reference computed
    (tests/golden/field_golden.json, g1_golden.json) and the EIP-196 vectors;
  - identity, doubling, inverse and scalar edge cases against textbook values."""
import ctypes
import json
import os
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P, R, G = M.BN_P, M.BN_R, M.CURVES[1]["g"]


@pytest.fixture(scope="module")
def lib():
    import crypto_amd

    lib = crypto_amd.lib.require_device()
    vp, u = ctypes.c_void_p, ctypes.c_uint
    for name in ("kinet_bn254_hip_g1_add", "kinet_bn254_hip_g1_mul", "kinet_bn254_hip_fp_mul"):
        getattr(lib, name).argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, u]
        getattr(lib, name).restype = ctypes.c_int
    assert lib.kinet_bn254_hip_available() == 1
    return lib


def _call(fn, a: bytes, b: bytes, out_rec: int, n: int) -> bytes:
    out = ctypes.create_string_buffer(out_rec * n)
    assert fn(a, b, out, n) == 0
    return out.raw


def _vectors(seed, n=100):
    rnd = random.Random(seed)
    pts = [M.mul(1, G, rnd.randrange(1, R)) for _ in range(n)]
    pts2 = [M.mul(1, G, rnd.randrange(1, R)) for _ in range(n)]
    pts2[:4] = [pts[0], M.neg(1, pts[1]), None, pts[3]]  # doubling, inverse, identity
    pts[3] = None
    ks = [rnd.getrandbits(256) for _ in range(n)]
    ks[:6] = [0, 1, R - 1, R, R + 1, (1 << 256) - 1]
    fa = [rnd.randrange(P) for _ in range(n)]
    fb = [rnd.randrange(P) for _ in range(n)]
    fa[:3], fb[:3] = [0, 1, P - 1], [P - 1, P - 1, P - 1]
    return pts, pts2, ks, fa, fb


def _enc_pts(pts):
    return b"".join(O.bn_point_ref(p) for p in pts)


def test_against_reference_driver_cpu_path(lib):
    if not O.bnref_available():
        pytest.skip("oracle/_ref not built")
    ref = O.bnref()
    n = 100
    pts, pts2, ks, fa, fb = _vectors(2540, n)
    a, b = _enc_pts(pts), _enc_pts(pts2)
    sc = b"".join(k.to_bytes(32, "little") for k in ks)
    ea, eb = b"".join(map(O.bn_fp_ref, fa)), b"".join(map(O.bn_fp_ref, fb))
    print("bn254 ops vs the reference driver's CPU path: 100 vectors each", flush=True)
    assert _call(lib.kinet_bn254_hip_g1_add, a, b, 72, n) == _call(ref.kinet_bn254_cuda_g1_add, a, b, 72, n)
    assert _call(lib.kinet_bn254_hip_g1_mul, a, sc, 72, n) == _call(ref.kinet_bn254_cuda_g1_mul, a, sc, 72, n)
    assert _call(lib.kinet_bn254_hip_fp_mul, ea, eb, 32, n) == _call(ref.kinet_bn254_cuda_fp_mul, ea, eb, 32, n)


def test_against_textbook_and_edge_cases(lib):
    n = 100
    pts, pts2, ks, fa, fb = _vectors(2541, n)
    print("bn254 ops vs textbook values", flush=True)
    got = _call(lib.kinet_bn254_hip_g1_add, _enc_pts(pts), _enc_pts(pts2), 72, n)
    assert [O.bn_point_from_ref(got[72 * i:72 * i + 72]) for i in range(n)] == \
        [M.add(1, x, y) for x, y in zip(pts, pts2)]
    assert got[72:144] == bytes(64) + (1).to_bytes(8, "little")  # P + (-P): x = y = 0, inf = 1
    sc = b"".join(k.to_bytes(32, "little") for k in ks)
    got = _call(lib.kinet_bn254_hip_g1_mul, _enc_pts(pts), sc, 72, n)
    assert [O.bn_point_from_ref(got[72 * i:72 * i + 72]) for i in range(n)] == \
        [M.mul(1, p, k) if p is not None else None for p, k in zip(pts, ks)]
    got = _call(lib.kinet_bn254_hip_fp_mul, b"".join(map(O.bn_fp_ref, fa)), b"".join(map(O.bn_fp_ref, fb)), 32, n)
    assert [O.bn_fp_from_ref(got[32 * i:32 * i + 32]) for i in range(n)] == [x * y % P for x, y in zip(fa, fb)]


def test_against_golden_and_eip196(lib):
    fcases = [c for c in json.load(open(os.path.join(GOLDEN, "field_golden.json"))) if c["field"] == O.FIELD_BN_FP]
    a = b"".join(O.bn_fp_ref(int.from_bytes(bytes.fromhex(c["a"]), "little")) for c in fcases)
    b = b"".join(O.bn_fp_ref(int.from_bytes(bytes.fromhex(c["b"]), "little")) for c in fcases)
    got = _call(lib.kinet_bn254_hip_fp_mul, a, b, 32, len(fcases))
    for i, c in enumerate(fcases):
        assert O.bn_fp_from_ref(got[32 * i:32 * i + 32]) == int.from_bytes(bytes.fromhex(c["mul"]), "little")
    gcases = [c for c in json.load(open(os.path.join(GOLDEN, "g1_golden.json"))) if c["curve"] == 1]

    def pt(hexs):
        w = bytes.fromhex(hexs)
        return M.from_wire(1, w)

    kG = [pt(c["kG"]) for c in gcases]
    k2G = [pt(c["k2G"]) for c in gcases]
    m = len(gcases)
    got = _call(lib.kinet_bn254_hip_g1_add, _enc_pts(kG), _enc_pts(k2G), 72, m)
    assert [O.bn_point_from_ref(got[72 * i:72 * i + 72]) for i in range(m)] == [pt(c["sum"]) for c in gcases]
    got = _call(lib.kinet_bn254_hip_g1_add, _enc_pts(kG), _enc_pts(kG), 72, m)
    assert [O.bn_point_from_ref(got[72 * i:72 * i + 72]) for i in range(m)] == [pt(c["dbl"]) for c in gcases]
    sc = b"".join(bytes.fromhex(c["k"]) for c in gcases)
    got = _call(lib.kinet_bn254_hip_g1_mul, _enc_pts([G] * m), sc, 72, m)
    assert [O.bn_point_from_ref(got[72 * i:72 * i + 72]) for i in range(m)] == kG
    e = json.load(open(os.path.join(GOLDEN, "bn254_eip196.json")))
    g = (int(e["G_x"], 16), int(e["G_y"], 16))
    two = (int(e["twoG_x"], 16), int(e["twoG_y"], 16))
    neg = (g[0], int(e["negG_y"], 16))
    got = _call(lib.kinet_bn254_hip_g1_add, _enc_pts([g, g]), _enc_pts([g, neg]), 72, 2)
    assert O.bn_point_from_ref(got[:72]) == two and O.bn_point_from_ref(got[72:]) is None
