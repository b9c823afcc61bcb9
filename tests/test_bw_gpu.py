"""Banderwagon (curve 3) MSMs on the GPU through gpukit_multi_pippenger_hip and
the prepared device path, against:
  - the reference's gpukit_multi_pippenger_cpu outputs (golden fixtures,
    status codes included, and the reference build live where oracle/_ref is
    present), 64-byte affine x || y;
  - the reference's 1000 MSM and 16 encoding known answers (gnark-crypto via
    the Go reference), through the device's own compressed output;
  - closed forms (sum s_i k_i mod r') G at 2^16 and 2^20 points.
Bar: bit-exact bytes."""
import hashlib
import json
import os
import random

import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLDEN, "msm_golden_bw.json")))
KATS = json.load(open(os.path.join(GOLDEN, "bw_kat.json")))
Q = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
R_BW = 0x1CFB69D4CA675F520CCE760202687600FF8F87007419047174FD06B52876E7E1
IDENTITY = bytes(32) + (1).to_bytes(32, "big")


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.lib.require_device()
    return crypto_amd


def _case_inputs(case):
    if "scalars" in case:
        return bytes.fromhex(case["scalars"]), bytes.fromhex(case["points"])
    rnd = random.Random(case["seed"])
    n = case["n"]
    sc = b"".join(rnd.randrange(R_BW).to_bytes(32, "little") for _ in range(n))
    ks = b"".join(rnd.getrandbits(253).to_bytes(32, "little") for _ in range(n))
    pts = O.bw_gen_points(ks)
    assert hashlib.sha256(pts).hexdigest() == case["sha256_points"]
    return sc, pts


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.get("name", f"n{c['n']}"))
def test_bw_golden(ca, case):
    sc, pts = _case_inputs(case)
    print(f"banderwagon golden n={case['n']} {case.get('name', '')}", flush=True)
    if case["status"] == 0:
        assert ca.multi_pippenger(3, sc, pts).hex() == case["expected"]
    else:
        with pytest.raises(ca.KinetError) as e:
            ca.multi_pippenger(3, sc, pts)
        assert e.value.code == case["status"] == -2


def _xorshift(seed):
    s = seed
    while True:
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        yield (s * 2685821657736338717) & 0xFFFFFFFFFFFFFFFF


def _kat_inputs(seed, n):
    """The input stream of banderwagon/test/multiexp_kat_test.cpp:72-110."""
    rng = _xorshift(seed)
    ks = b"".join(((next(rng) & 0xFFFFFFFF) + 1).to_bytes(32, "little") for _ in range(n))
    sc = bytearray()
    for _ in range(n):
        buf = b"".join(next(rng).to_bytes(8, "little") for _ in range(4))
        sc += buf[:31] + bytes([buf[31] & 0x07])
    return bytes(sc), O.bw_gen_points(ks)


def test_bw_msm_kats(ca):
    """All 1000 known answers: the device's compressed output (prepared path)
    and the host entry's affine output, compressed by the oracle."""
    import torch
    from crypto_amd import device as D

    bad = []
    for i, kat in enumerate(KATS["msm"]):
        sc, pts = _kat_inputs(kat["seed"], kat["n"])
        if i % 100 == 0:
            print(f"banderwagon kat {kat['name']} n={kat['n']}", flush=True)
        wire = ca.multi_pippenger(3, sc, pts)
        pp = D.PreparedPoints(3, pts)
        rec = D.msm_prepared(pp, D.to_device(sc)).cpu().numpy().tobytes()
        if O.bw_compress(wire).hex() != kat["enc_be"] or rec[:64] != wire or rec[64:96].hex() != kat["enc_be"]:
            bad.append(kat["name"])
    torch.cuda.synchronize()
    assert not bad, bad[:10]


def _fixed_base(ca, ks: bytes):
    import torch
    from crypto_amd import device as D

    n = len(ks) // 32
    out = torch.empty(n * 64, dtype=torch.uint8, device="cuda")
    rc = ca.library().kinet_hip_g1_fixed_base_mul(3, D.to_device(ks).data_ptr(), n, out.data_ptr(),
                                                  D.stream_handle())
    assert rc == 0, rc
    torch.cuda.synchronize()
    return out


def test_bw_fixed_base_encoding_kats(ca):
    """k G on the device for the 16 encoding known answers (their compressed
    forms) and random k (the oracle's multiple)."""
    ks = [bytes.fromhex(k["scalar_le"]) for k in KATS["enc"]]
    rnd = random.Random(5)
    extra = [rnd.randrange(R_BW) for _ in range(16)] + [R_BW, R_BW + 1, 0]
    kb = b"".join(ks) + b"".join(k.to_bytes(32, "little") for k in extra)
    out = _fixed_base(ca, kb).cpu().numpy().tobytes()
    for i, kat in enumerate(KATS["enc"]):
        assert O.bw_compress(out[64 * i: 64 * i + 64]).hex() == kat["enc_be"]
    for j, k in enumerate(extra):
        got = out[64 * (16 + j): 64 * (17 + j)]
        want = O.bw_mul_gen(k)
        assert got == (bytes(64) if want == IDENTITY else want)  # the identity leaves as zero bytes


@pytest.mark.parametrize("log2n", [16, 20])
def test_bw_closed_form(ca, log2n):
    """sum s_i (k_i G) = (sum s_i k_i mod r') G at 2^16 and 2^20 points, with
    edge scalars 0, 1 and r' - 1 and identity points in the first slots."""
    import torch
    from crypto_amd import device as D

    n = 1 << log2n
    rnd = random.Random(log2n)
    ks = [rnd.randrange(R_BW) for _ in range(n)]
    ss = [rnd.randrange(R_BW) for _ in range(n)]
    ss[:3] = [0, 1, R_BW - 1]
    ks[3] = 0  # an identity point (zero bytes from the generator)
    kb = b"".join(k.to_bytes(32, "little") for k in ks)
    sb = b"".join(s.to_bytes(32, "little") for s in ss)
    print(f"banderwagon closed form n=2^{log2n}", flush=True)
    pp = D.PreparedPoints(3, _fixed_base(ca, kb))
    rec = D.msm_prepared(pp, D.to_device(sb)).cpu().numpy().tobytes()
    t = sum(s * k for s, k in zip(ss, ks)) % R_BW
    want = O.bw_mul_gen(t)
    assert rec[:64] == want
    assert rec[64:96] == O.bw_compress(want)
    torch.cuda.synchronize()


def test_bw_every_window_size(ca):
    from crypto_amd import device as D

    rnd = random.Random(9)
    n = 3000
    ks = [rnd.randrange(R_BW) for _ in range(n)]
    ss = [rnd.randrange(R_BW) for _ in range(n)]
    pp = D.PreparedPoints(3, _fixed_base(ca, b"".join(k.to_bytes(32, "little") for k in ks)))
    d_sc = D.to_device(b"".join(s.to_bytes(32, "little") for s in ss))
    want = O.bw_mul_gen(sum(s * k for s, k in zip(ss, ks)) % R_BW)
    for c in range(5, 19):
        rec = D.msm_prepared(pp, d_sc, window_bits=c).cpu().numpy().tobytes()
        assert rec[:64] == want, c


def test_bw_empty_and_refusals(ca):
    import torch
    from crypto_amd import device as D

    lib = ca.library()
    assert ca.multi_pippenger(3, b"", b"") == IDENTITY
    pp = D.PreparedPoints(3, O.bw_gen_points((5).to_bytes(32, "little")))
    rec = D.msm_prepared(pp, D.to_device(bytes(32)), n=0).cpu().numpy().tobytes()
    assert rec[:96] == IDENTITY + bytes(32)
    # an invalid point: prepare refuses it
    bad = O.bw_generator()[:32] + (5).to_bytes(32, "big")
    d_bad = D.to_device(bad)
    pts = torch.empty(256, dtype=torch.uint8, device="cuda")
    inf = torch.empty(16, dtype=torch.uint8, device="cuda")
    assert lib.kinet_hip_points_prepare(3, d_bad.data_ptr(), 1, pts.data_ptr(), inf.data_ptr(),
                                        D.stream_handle()) == -2
    # the G1 (GLV) form, window tables and the no-allocation entry are Weierstrass only
    assert lib.kinet_hip_msm_prepared_g1(3, None, None, None, 0, 0, 0, None, None) == -3
    assert lib.kinet_hip_table_bytes(3, 16, 8) == 0
    assert lib.kinet_hip_msm_prepared_noalloc(3, None, None, None, 0, None, None) == -3


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent (make -C oracle ref)")
@pytest.mark.parametrize("n", [2, 3, 97, 1500])
def test_bw_matches_reference_live(ca, n):
    """The reference's own CPU path on random inputs, including the other
    representative (-x, -y) of some points and the two-torsion point (0, -1)."""
    rnd = random.Random(1000 + n)
    pts = bytearray(O.bw_gen_points(b"".join(rnd.getrandbits(253).to_bytes(32, "little") for _ in range(n))))
    for i in range(0, n, 4):
        x, y = int.from_bytes(pts[64 * i: 64 * i + 32], "big"), int.from_bytes(pts[64 * i + 32: 64 * i + 64], "big")
        pts[64 * i: 64 * i + 64] = ((Q - x) % Q).to_bytes(32, "big") + ((Q - y) % Q).to_bytes(32, "big")
    if n > 2:
        pts[64:128] = bytes(32) + (Q - 1).to_bytes(32, "big")
    sc = b"".join(rnd.randrange(R_BW).to_bytes(32, "little") for _ in range(n))
    rc, want = O.ref_bw_msm(sc, bytes(pts))
    assert rc == 0
    assert ca.multi_pippenger(3, sc, bytes(pts)) == want
