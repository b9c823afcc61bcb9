"""GPU parity of the HIP MSM against the oracle and the reference's golden
outputs, through the C ABI.  Bar: bit-exact wire bytes."""
import hashlib
import json
import os
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MSM_CASES = json.load(open(os.path.join(GOLDEN, "msm_golden.json")))


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.lib.require_device()
    return crypto_amd


def _inputs(case):
    if "scalars" in case:
        return bytes.fromhex(case["scalars"]), bytes.fromhex(case["points"])
    curve, n = case["curve"], case["n"]
    scalars = O.gen_scalars(curve, case["seed"], n)
    points = O.gen_points(curve, O.gen_scalars(curve, case["seed"] ^ 0x5EED, n))
    assert hashlib.sha256(points).hexdigest() == case["sha256_points"]
    return scalars, points


@pytest.mark.parametrize("case", MSM_CASES, ids=lambda c: f"c{c['curve']}-n{c['n']}-{c.get('name', 'seeded')}")
def test_msm_golden(ca, case):
    scalars, points = _inputs(case)
    print(f"golden curve={case['curve']} n={case['n']} {case.get('name', '')}", flush=True)
    assert ca.multi_pippenger(case["curve"], scalars, points).hex() == case["expected"]


@pytest.mark.parametrize("curve", [1, 2])
def test_msm_empty_and_backend_env(ca, curve, monkeypatch):
    """n = 0 is the identity; the HIP backend itself does not read
    GPUKIT_BACKEND (backend selection is the reference dispatcher's)."""
    assert ca.multi_pippenger(curve, b"", b"") == bytes(ca.point_bytes(curve))
    g = O.generator(curve)
    for env in ("cpu", "hip"):
        monkeypatch.setenv("GPUKIT_BACKEND", env)
        assert ca.multi_pippenger(curve, (2).to_bytes(32, "little"), g) == O.g1_dbl(curve, g)


def test_msm_unsupported_curves(ca):
    with pytest.raises(ca.KinetError) as e:  # no curve 4 (GPUKIT_ERR_BACKEND, the reference's code)
        ca.multi_pippenger(4, bytes(32), bytes(64))
    assert e.value.code == -3


def _closed_form(curve, scalars, k0, d):
    r = M.CURVES[curve]["r"]
    n = len(scalars) // 32
    t = 0
    for i in range(n):
        t += int.from_bytes(scalars[32 * i:32 * i + 32], "little") * ((k0 + i * d) % r)
    return O.g1_mul_gen(curve, (t % r).to_bytes(32, "little"))


@pytest.mark.parametrize("curve,n", [(1, 1 << 16), (2, 1 << 16), (1, 100003)])
def test_msm_large_closed_form(ca, curve, n):
    k0, d = 0x1234567890ABCDEF1234567, 0xFEDCBA987654321
    points = O.gen_points_progression(curve, k0, d, n)
    scalars = O.gen_scalars(curve, 4242 + n, n)
    print(f"closed form curve={curve} n={n}", flush=True)
    assert ca.multi_pippenger(curve, scalars, points) == _closed_form(curve, scalars, k0, d)


@pytest.mark.parametrize("curve", [1, 2])
def test_msm_every_window_size(ca, curve):
    """All instantiated window sizes (incl. c = 8, 16 with the carry window)
    must give the same bytes, also for scalars >= 2^255.  n = 5000 takes the
    LDS-histogram sort for c <= 16 and the global-atomic sort for c = 17, 18."""
    import torch
    from crypto_amd import device as D

    n = 5000
    k0, d = 77, 0x9999999999999999
    points = O.gen_points_progression(curve, k0, d, n)
    rnd = random.Random(curve)
    sc = [rnd.getrandbits(256) for _ in range(n)]
    sc[:50] = [(1 << 256) - 1 - i for i in range(50)]       # top carry in every window
    sc[50:60] = [0] * 10
    scalars = b"".join(s.to_bytes(32, "little") for s in sc)
    want = _closed_form(curve, scalars, k0, d)
    pp = D.PreparedPoints(curve, points)
    d_sc = D.to_device(scalars)
    pb = ca.point_bytes(curve)
    for c in range(5, 19):
        print(f"window bits c={c} curve={curve}", flush=True)
        out = D.msm_prepared(pp, d_sc, window_bits=c).cpu().numpy().tobytes()
        assert out[:pb] == want, c
        assert out[pb:] == O.compress(curve, want), c
    del torch


@pytest.mark.parametrize("curve", [1, 2])
@pytest.mark.parametrize("c", [0, 7, 16])
def test_msm_sort_paths_agree_with_identity_points(ca, curve, c):
    """Identity points (all-zero wire) and zero scalars inside an MSM large
    enough for the LDS sort; the same MSM with c = 17 (global-atomic sort)
    and the closed form over the non-identity points must agree."""
    from crypto_amd import device as D

    n = 20000
    k0, d = 4321, 0x77777777
    pb = O.POINT_BYTES[curve]
    points = bytearray(O.gen_points_progression(curve, k0, d, n))
    zero_at = set(range(0, n, 97))
    for i in zero_at:
        points[i * pb:(i + 1) * pb] = bytes(pb)
    rnd = random.Random(31 + c)
    sc = [0 if i % 89 == 0 else rnd.getrandbits(255) for i in range(n)]
    scalars = b"".join(s.to_bytes(32, "little") for s in sc)
    r = M.CURVES[curve]["r"]
    t = sum(s * (k0 + i * d) for i, s in enumerate(sc) if i not in zero_at) % r
    want = O.g1_mul_gen(curve, t.to_bytes(32, "little"))
    pp = D.PreparedPoints(curve, bytes(points))
    d_sc = D.to_device(scalars)
    print(f"sort paths curve={curve} c={c}", flush=True)
    got = D.msm_prepared(pp, d_sc, window_bits=c).cpu().numpy().tobytes()
    alt = D.msm_prepared(pp, d_sc, window_bits=17).cpu().numpy().tobytes()
    assert got[:pb] == want
    assert alt == got


@pytest.mark.parametrize("curve", [1, 2])
def test_msm_adversarial_buckets(ca, curve):
    """Skewed inputs: every scalar equal (one bucket per window holds all
    points, spanning many accumulation chunks), repeated points (doubling
    inside buckets) and cancelling pairs."""
    n = 5000
    g = O.generator(curve)
    pts = O.gen_points_progression(curve, 5, 3, 64)
    pb = O.POINT_BYTES[curve]
    same = (0xABCDEF0123456789 << 180).to_bytes(32, "little")
    scalars = same * n
    points = b"".join(pts[pb * (i % 64): pb * (i % 64 + 1)] for i in range(n))
    want = O.msm(curve, scalars[:64 * 32], points[:64 * pb])  # n = 5000 = 78*64 + 8
    got = ca.multi_pippenger(curve, scalars, points)
    r = M.CURVES[curve]["r"]
    k = sum((5 + 3 * (i % 64)) for i in range(n)) * int.from_bytes(same, "little") % r
    assert got == O.g1_mul_gen(curve, k.to_bytes(32, "little"))
    assert want  # oracle sanity
    negg = O.g1_mul_gen(curve, (r - 1).to_bytes(32, "little"))
    sc = O.gen_scalars(curve, 9, 1)
    assert ca.multi_pippenger(curve, sc * 2, g + negg) == bytes(pb)
    assert ca.multi_pippenger(curve, sc * 4, g * 4) == O.g1_mul_gen(
        curve, (4 * int.from_bytes(sc, "little") % r).to_bytes(32, "little"))


@pytest.mark.parametrize("curve", [1, 2])
def test_fixed_base_generator(ca, curve):
    """kinet_hip_g1_fixed_base_mul against the oracle's k * G."""
    import torch
    from crypto_amd import device as D

    r = M.CURVES[curve]["r"]
    rnd = random.Random(17 + curve)
    ks = [0, 1, 2, 255, 256, r - 1, r, r + 1, (1 << 256) - 1] + [rnd.getrandbits(256) for _ in range(40)]
    kb = b"".join(k.to_bytes(32, "little") for k in ks)
    pb = ca.point_bytes(curve)
    out = torch.empty(pb * len(ks), dtype=torch.uint8, device="cuda")
    lib = ca.library()
    assert lib.kinet_hip_g1_fixed_base_mul(curve, D.to_device(kb).data_ptr(), len(ks), out.data_ptr(),
                                           D.stream_handle()) == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy().tobytes()
    for i, k in enumerate(ks):
        assert got[pb * i: pb * (i + 1)] == O.g1_mul_gen(curve, k.to_bytes(32, "little")), (i, k)


def test_msm_bn254_host_entry_off_curve_point(ca):
    """gpukit_multi_pippenger_hip runs BN254 inputs whose points all lie on the
    curve through the GLV split (cofactor 1; test_msm_large_closed_form checks
    those against the closed form) and any other input through the plain path.
    With one point off the curve, carrying scalar 0, the plain path's result is
    the closed form over the other points.  (Off-curve points with nonzero
    scalars have no defined sum: the reference reads coordinates mod p without
    checking them, and the addition formulas are not a group law off the
    curve.)"""
    n = 5000
    k0, d = 0x1234567, 0x89ABCDEF
    points = bytearray(O.gen_points_progression(1, k0, d, n))
    y = int.from_bytes(points[64 * 17 + 32: 64 * 18], "big")
    points[64 * 17 + 32: 64 * 18] = ((y + 1) % M.CURVES[1]["p"]).to_bytes(32, "big")
    assert not O.on_curve(1, bytes(points[64 * 17: 64 * 18]))
    scalars = bytearray(O.gen_scalars(1, 77, n))
    scalars[32 * 17: 32 * 18] = bytes(32)
    scalars = bytes(scalars)
    assert ca.multi_pippenger(1, scalars, bytes(points)) == _closed_form(1, scalars, k0, d)


def _non_g1_point(curve):
    """A point of E(Fp) outside G1 (BLS12-381 has a large cofactor)."""
    p, b = M.CURVES[curve]["p"], M.CURVES[curve]["b"]
    for x in range(1, 1000):
        y = M.sqrt(curve, (x * x * x + b) % p)
        if y is not None:
            P = (x, y)
            if M.mul(curve, P, M.CURVES[curve]["r"]) is not None:
                return P
    raise AssertionError("no point found")


@pytest.mark.parametrize("curve", [1, 2])
def test_check_g1(ca, curve):
    """kinet_hip_points_check_g1: multiples of G pass (identity points are
    skipped); a point off the curve fails; for BLS12-381 a point of E(Fp)
    outside G1 fails."""
    from crypto_amd import device as D

    pb = O.POINT_BYTES[curve]
    good = O.gen_points_progression(curve, 11, 13, 300) + bytes(pb)
    assert D.check_g1(D.PreparedPoints(curve, good))
    off = bytearray(good)
    off[pb - 1] ^= 1  # flip the low bit of the first point's y
    assert not D.check_g1(D.PreparedPoints(curve, bytes(off)))
    if curve == 2:
        P = _non_g1_point(curve)
        assert M.on_curve(curve, P)
        bad = good[: pb * 10] + M.to_wire(curve, P) + good[pb * 10:]
        assert not D.check_g1(D.PreparedPoints(curve, bad))


@pytest.mark.parametrize("n", [2048, 5000, 1 << 16])
@pytest.mark.parametrize("c", [0, 9, 12, 15, 16])
def test_msm_g1_glv_matches_plain(ca, n, c):
    """BLS12-381 through the GLV split (kinet_hip_msm_prepared_g1) gives the
    plain path's bytes and the closed form, for scalars of every range
    (0, r - 1, r, >= 2^255, 2^256 - 1, lambda) and identity points."""
    from crypto_amd import device as D

    if n == 1 << 16 and c not in (0, 16):
        pytest.skip("large size covered at the default and widest window")
    curve = 2
    k0, d = 0x5A5A5A5A, 0x1F2E3D4C5B6A7
    pb = O.POINT_BYTES[curve]
    points = bytearray(O.gen_points_progression(curve, k0, d, n))
    zero_at = set(range(3, n, 211))
    for i in zero_at:
        points[i * pb:(i + 1) * pb] = bytes(pb)
    r = M.CURVES[curve]["r"]
    lam = 0xAC45A4010001A40200000000FFFFFFFF
    rnd = random.Random(n + c)
    sc = [rnd.getrandbits(256) for _ in range(n)]
    sc[:8] = [0, 1, r - 1, r, r + 1, (1 << 256) - 1, lam, lam * lam % r]
    sc[8:20] = [(1 << 255) + rnd.getrandbits(200) for _ in range(12)]
    scalars = b"".join(s.to_bytes(32, "little") for s in sc)
    t = sum(s * (k0 + i * d) for i, s in enumerate(sc) if i not in zero_at) % r
    want = O.g1_mul_gen(curve, t.to_bytes(32, "little"))
    pp = D.PreparedPoints(curve, bytes(points))
    d_sc = D.to_device(scalars)
    print(f"glv n={n} c={c}", flush=True)
    got = D.msm_prepared(pp, d_sc, window_bits=c, g1=True).cpu().numpy().tobytes()
    plain = D.msm_prepared(pp, d_sc, window_bits=(c if c else 0)).cpu().numpy().tobytes()
    assert got[:pb] == want
    assert got == plain


def test_msm_g1_golden_bls(ca):
    """The reference's golden BLS12-381 outputs through the GLV path."""
    from crypto_amd import device as D

    for case in MSM_CASES:
        if case["curve"] != 2 or case["n"] < 2048 or "scalars" in case:
            continue
        scalars, points = _inputs(case)
        pp = D.PreparedPoints(2, points)
        assert D.check_g1(pp)
        out = D.msm_prepared(pp, D.to_device(scalars), g1=True).cpu().numpy().tobytes()
        assert out[:96].hex() == case["expected"], case["n"]


@pytest.mark.parametrize("n,c", [(2048, 0), (5000, 0), (5000, 9), (5000, 16), (1 << 16, 0)])
def test_msm_g1_glv_bn254(ca, n, c):
    """BN254 through the signed GLV split (kinet_hip_msm_prepared_g1): the plain
    path's bytes and the closed form, for scalars 0, 1, r - 1, r, r + 1,
    2^256 - 1, lambda, lambda^2, values near multiples of r and random ones,
    with identity points among the inputs."""
    from crypto_amd import device as D

    curve = 1
    k0, d = 0x31337, 0xBEEF1234567
    pb = O.POINT_BYTES[curve]
    points = bytearray(O.gen_points_progression(curve, k0, d, n))
    zero_at = set(range(5, n, 197))
    for i in zero_at:
        points[i * pb:(i + 1) * pb] = bytes(pb)
    r = M.CURVES[curve]["r"]
    lam = 0x30644E72E131A029048B6E193FD84104CC37A73FEC2BC5E9B8CA0B2D36636F23
    rnd = random.Random(n * 31 + c)
    sc = [rnd.getrandbits(256) for _ in range(n)]
    sc[:12] = [0, 1, r - 1, r, r + 1, (1 << 256) - 1, lam, lam * lam % r, 2 * r - 1, 4 * r + 7, r - lam, (r + 1) // 2]
    scalars = b"".join(s.to_bytes(32, "little") for s in sc)
    t = sum(s * (k0 + i * d) for i, s in enumerate(sc) if i not in zero_at) % r
    want = O.g1_mul_gen(curve, t.to_bytes(32, "little"))
    pp = D.PreparedPoints(curve, bytes(points))
    assert D.check_g1(pp)
    d_sc = D.to_device(scalars)
    print(f"bn glv n={n} c={c}", flush=True)
    got = D.msm_prepared(pp, d_sc, window_bits=c, g1=True).cpu().numpy().tobytes()
    plain = D.msm_prepared(pp, d_sc).cpu().numpy().tobytes()
    assert got[:pb] == want
    assert got == plain


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_points_prepare_fills_exactly_the_queried_size(ca, curve):
    """kinet_hip_points_prepare writes kinet_hip_points_bytes(curve, n) bytes of
    records and n flag bytes, nothing beyond: guard bands after both buffers
    keep their fill pattern, and the records give the closed-form MSM."""
    import torch
    from crypto_amd import device as D

    lib = ca.library()
    n = 777
    want_rec = {0: 80, 1: 80, 2: 128}[curve]
    nbytes = lib.kinet_hip_points_bytes(curve, n)
    assert nbytes == want_rec * n
    assert lib.kinet_hip_points_bytes(3, n) == 128 * n
    assert lib.kinet_hip_points_bytes(4, n) == 0
    k0, d = 99, 0x10001
    if curve == 0:  # secp256k1: the oracle has no secp256k1; pure-Python points
        G, r = M.CURVES[0]["g"], M.CURVES[0]["r"]
        P, Dd, pl = M.mul(0, G, k0), M.mul(0, G, d), []
        for _ in range(n):
            pl.append(M.to_wire(0, P))
            P = M.add(0, P, Dd)
        points = b"".join(pl)
        rnd = random.Random(5150)
        scalars = b"".join(rnd.getrandbits(255).to_bytes(32, "little") for _ in range(n))
        t = sum(int.from_bytes(scalars[32 * i:32 * i + 32], "little") * (k0 + i * d) for i in range(n)) % r
        want = M.to_wire(0, M.mul(0, G, t)) if t else bytes(64)
    else:
        points = O.gen_points_progression(curve, k0, d, n)
        scalars = O.gen_scalars(curve, 5150, n)
        want = _closed_form(curve, scalars, k0, d)
    guard = 4096
    pts = torch.full((nbytes + guard,), 0xA5, dtype=torch.uint8, device="cuda")
    inf = torch.full((n + guard,), 0x5A, dtype=torch.uint8, device="cuda")
    d_wire = D.to_device(points)
    assert lib.kinet_hip_points_prepare(curve, d_wire.data_ptr(), n, pts.data_ptr(), inf.data_ptr(),
                                        D.stream_handle()) == 0
    torch.cuda.synchronize()
    assert bool((pts[nbytes:] == 0xA5).all()), "points_prepare wrote past kinet_hip_points_bytes"
    assert bool((inf[n:] == 0x5A).all()), "points_prepare wrote past n flag bytes"
    out = torch.empty(ca.point_bytes(curve) * 2, dtype=torch.uint8, device="cuda")
    assert lib.kinet_hip_msm_prepared(curve, D.to_device(scalars).data_ptr(), pts.data_ptr(), inf.data_ptr(), n, 0,
                                      0, out.data_ptr(), D.stream_handle()) == 0
    torch.cuda.synchronize()
    pb = ca.point_bytes(curve)
    assert out.cpu().numpy().tobytes()[:pb] == want


@pytest.mark.parametrize("g1", [False, True])
def test_shared_workspace_ordered_across_streams(ca, g1):
    """Back-to-back kinet_hip_msm_prepared(_g1) calls on two different streams
    share the library's MSM workspace: the second stream waits for the first
    MSM, so both results are right (a large MSM first, then a smaller one)."""
    import torch
    from crypto_amd import device as D

    curve = 2
    n1, n2 = 1 << 17, 5000
    k0, d = 0x1234, 0x777
    points = O.gen_points_progression(curve, k0, d, n1)
    pp = D.PreparedPoints(curve, points)
    sc1 = O.gen_scalars(curve, 61, n1)
    sc2 = O.gen_scalars(curve, 62, n2)
    d1, d2 = D.to_device(sc1), D.to_device(sc2)
    want1 = _closed_form(curve, sc1, k0, d)
    want2 = _closed_form(curve, sc2, k0, d)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for rep in range(2):
        print(f"two streams g1={g1} rep={rep}", flush=True)
        o1 = D.msm_prepared(pp, d1, stream=sa, sync=False, g1=g1)
        o2 = D.msm_prepared(pp, d2, n=n2, stream=sb, sync=False, g1=g1)
        torch.cuda.synchronize()
        assert o1.cpu().numpy().tobytes()[:96] == want1
        assert o2.cpu().numpy().tobytes()[:96] == want2
