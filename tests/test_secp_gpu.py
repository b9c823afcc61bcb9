"""secp256k1 (curve 0) MSMs through gpukit_multi_pippenger_hip and the prepared
plain path: the reference's own outputs (golden fixtures, and the compiled
reference CPU Pippenger live where oracle/_ref is present), closed forms
sum s_i (k0 + i d) G for full-range scalars, every window size, edge cases
and refusals."""
import hashlib
import json
import os
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLDEN, "msm_golden_secp.json")))
G, N = M.CURVES[0]["g"], M.CURVES[0]["r"]


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.lib.require_device()
    return crypto_amd


def _points(n, k0, d):
    P, D = M.mul(0, G, k0), M.mul(0, G, d)
    out = []
    for _ in range(n):
        out.append(M.to_wire(0, P))
        P = M.add(0, P, D)
    return b"".join(out)


def _closed_form(scalars, k0, d):
    t = sum(int.from_bytes(scalars[32 * i:32 * i + 32], "little") * ((k0 + i * d) % N)
            for i in range(len(scalars) // 32))
    return M.to_wire(0, M.mul(0, G, t % N)) if t % N else bytes(64)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c['n']}-{c.get('name', 'seeded')}")
def test_secp_msm_golden(ca, case):
    if "scalars" in case:
        scalars, points = bytes.fromhex(case["scalars"]), bytes.fromhex(case["points"])
    else:
        rnd = random.Random(case["seed"])
        scalars = b"".join(rnd.getrandbits(255).to_bytes(32, "little") for _ in range(case["n"]))
        points = _points(case["n"], case["k0"], case["d"])
        assert hashlib.sha256(points).hexdigest() == case["sha256_points"]
    print(f"secp256k1 golden n={case['n']} {case.get('name', '')}", flush=True)
    assert ca.multi_pippenger(0, scalars, points).hex() == case["expected"]


@pytest.mark.parametrize("n", [3, 777, 5000])
def test_secp_msm_matches_reference_live(ca, n):
    """Random points and scalars below 2^255 against the reference's compiled
    CPU Pippenger (oracle/_ref), with identity points and repeats mixed in."""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    rnd = random.Random(n)
    pts = _points(n, rnd.getrandbits(250), rnd.getrandbits(250))
    pl = [pts[64 * i:64 * i + 64] for i in range(n)]
    for i in range(0, n, 17):
        pl[i] = bytes(64)
    for i in range(5, n, 23):
        pl[i] = pl[i - 1]
    points = b"".join(pl)
    scalars = b"".join(rnd.getrandbits(255).to_bytes(32, "little") for _ in range(n))
    print(f"secp256k1 vs reference n={n}", flush=True)
    assert ca.multi_pippenger(0, scalars, points) == O.ref_msm(0, scalars, points)


@pytest.mark.parametrize("n", [1000, 1 << 16])
def test_secp_msm_closed_form_full_range(ca, n):
    """Scalars over all of [0, 2^256), with n - 1, n, 2^256 - 1 and zeros:
    the result is sum s_i P_i (the final window carry is kept)."""
    k0, d = 0xC0FFEE12345, 0xBADC0DE987654321
    points = _points(n, k0, d)
    rnd = random.Random(7 + n)
    sc = [rnd.getrandbits(256) for _ in range(n)]
    sc[:6] = [N - 1, N, N + 1, (1 << 256) - 1, 0, 1 << 255]
    scalars = b"".join(s.to_bytes(32, "little") for s in sc)
    print(f"secp256k1 closed form n={n}", flush=True)
    assert ca.multi_pippenger(0, scalars, points) == _closed_form(scalars, k0, d)


def test_secp_msm_every_window_size(ca):
    """Every instantiated window size of the prepared plain path gives the
    same bytes (c = 8, 16 with the carry window; 17, 18 with the global sort)."""
    from crypto_amd import device as D

    n, k0, d = 5000, 91, 0x7777777777777777
    points = _points(n, k0, d)
    rnd = random.Random(3)
    sc = [rnd.getrandbits(256) for _ in range(n)]
    sc[:40] = [(1 << 256) - 1 - i for i in range(40)]
    sc[40:50] = [0] * 10
    scalars = b"".join(s.to_bytes(32, "little") for s in sc)
    want = _closed_form(scalars, k0, d)
    pp = D.PreparedPoints(0, points)
    d_sc = D.to_device(scalars)
    for c in range(5, 19):
        print(f"secp256k1 window bits c={c}", flush=True)
        out = D.msm_prepared(pp, d_sc, window_bits=c).cpu().numpy().tobytes()
        assert out[:64] == want, c
        assert out[64:] == bytes(32), c  # no compressed form for secp256k1


def test_secp_msm_empty_and_refusals(ca):
    assert ca.multi_pippenger(0, b"", b"") == bytes(64)
    assert ca.multi_pippenger(0, (3).to_bytes(32, "little"), M.to_wire(0, G)) == M.to_wire(0, M.mul(0, G, 3))
    lib = ca.lib.require_device()
    # window tables are BN254 / BLS12-381 only; the G1 (GLV) form takes
    # secp256k1 (a missing output is then a null argument, not a bad curve)
    assert lib.kinet_hip_table_bytes(0, 16, 8) == 0
    assert lib.kinet_hip_msm_prepared_g1(0, None, None, None, 0, 0, 0, None, None) == -1
    assert lib.kinet_hip_msm_prepared_g1(4, None, None, None, 0, 0, 0, None, None) == -3
    with pytest.raises(ca.KinetError) as e:
        ca.multi_pippenger(4, bytes(32), bytes(64))
    assert e.value.code == -3


def test_secp_fixed_base(ca):
    """kinet_hip_g1_fixed_base_mul on curve 0 (the bench's point generator)
    against double-and-add, for 0, 1, n - 1, n, 2^256 - 1 and random scalars."""
    import torch

    lib = ca.lib.require_device()
    rnd = random.Random(11)
    ks = [0, 1, 2, N - 1, N, (1 << 256) - 1] + [rnd.getrandbits(256) for _ in range(40)]
    d_k = torch.tensor(list(b"".join(k.to_bytes(32, "little") for k in ks)), dtype=torch.uint8, device="cuda")
    out = torch.empty(64 * len(ks), dtype=torch.uint8, device="cuda")
    assert lib.kinet_hip_g1_fixed_base_mul(0, d_k.data_ptr(), len(ks), out.data_ptr(), None) == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy().tobytes()
    for i, k in enumerate(ks):
        want = M.to_wire(0, M.mul(0, G, k % N)) if k % N else bytes(64)
        assert got[64 * i:64 * i + 64] == want, k


LAMBDA = 0x5363AD4CC05C30E0A5261C028812645A122E22EA20816678DF02967C1B23BD72


def _glv_inputs(n, k0, d, seed):
    """Progression points with identities every 211th, and scalars 0, 1,
    n - 1, n, n + 1, 2^256 - 1, lambda, lambda^2, -lambda, 2 lambda,
    2^256 - n, (n + 1) / 2 and random ones (about a fifth of the random ones have a
    GLV half of 2^127 or more, so the 128th bit is exercised)."""
    pl = bytearray(_points(n, k0, d))
    zero_at = set(range(3, n, 211))
    for i in zero_at:
        pl[64 * i:64 * i + 64] = bytes(64)
    rnd = random.Random(seed)
    sc = [rnd.getrandbits(256) for _ in range(n)]
    sc[:12] = [0, 1, N - 1, N, N + 1, (1 << 256) - 1, LAMBDA, LAMBDA * LAMBDA % N, N - LAMBDA, 2 * LAMBDA,
               (1 << 256) - N, (N + 1) // 2]
    scalars = b"".join(x.to_bytes(32, "little") for x in sc)
    t = sum(x * ((k0 + i * d) % N) for i, x in enumerate(sc) if i not in zero_at) % N
    want = M.to_wire(0, M.mul(0, G, t)) if t else bytes(64)
    return bytes(pl), scalars, want


@pytest.mark.parametrize("n,c", [(2048, 0), (5000, 0), (5000, 8), (5000, 9), (5000, 16), (1 << 16, 0),
                                 (1 << 16, 13)])
def test_secp_msm_glv(ca, n, c):
    """secp256k1 through the signed GLV split (kinet_hip_msm_prepared_g1,
    phi(x, y) = (beta x, y) = lambda (x, y), halves below 2^128 in 128-bit
    windows; c = 8 and 16 take the carry window): the plain path's bytes and
    the closed form sum s_i (k0 + i d) G."""
    from crypto_amd import device as D

    k0, d = 0x5EC9256B1, 0xF00DFACE12345
    points, scalars, want = _glv_inputs(n, k0, d, n * 7 + c)
    pp = D.PreparedPoints(0, points)
    assert D.check_g1(pp)
    d_sc = D.to_device(scalars)
    print(f"secp256k1 glv n={n} c={c}", flush=True)
    got = D.msm_prepared(pp, d_sc, window_bits=c, g1=True).cpu().numpy().tobytes()
    plain = D.msm_prepared(pp, d_sc).cpu().numpy().tobytes()
    assert got[:64] == want
    assert got == plain


def test_secp_host_entry_glv_and_off_curve(ca):
    """gpukit_multi_pippenger_hip on secp256k1 at n >= 4096 checks the points
    and runs the GLV split when all lie on the curve (cofactor 1): the closed
    form.  With one point off the curve (scalar 0) it keeps the plain path and
    still gives the closed form; check_g1 flags that point."""
    from crypto_amd import device as D

    n, k0, d = 6000, 0x1234567, 0xABCDEF0123
    points, scalars, want = _glv_inputs(n, k0, d, 99)
    print("secp256k1 host entry glv", flush=True)
    assert ca.multi_pippenger(0, scalars, points) == want
    bad = bytearray(points)
    bad[64 * 3 + 63] ^= 1  # point 3 (an identity) becomes (0, 1), off the curve
    sc = bytearray(scalars)
    sc[32 * 3:32 * 4] = bytes(32)
    assert not D.check_g1(D.PreparedPoints(0, bytes(bad)))
    print("secp256k1 host entry off-curve point", flush=True)
    assert ca.multi_pippenger(0, bytes(sc), bytes(bad)) == want
