"""GPU parity of the window-table MSM (kinet_hip_points_table_build +
kinet_hip_msm_table): equal bytes to the plain MSM and to the closed form over
points in arithmetic progression, for both sort paths, several window widths,
identity points, prefixes of the table and one bucket holding every digit."""
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    import crypto_amd
    from crypto_amd import device as D

    crypto_amd.lib.require_device()
    return D


def _case(curve, n, seed, zero_every=0):
    k0, d = 0x7777 + seed, 0x1234567890ABCDEF + 5 * seed
    pb = O.POINT_BYTES[curve]
    points = bytearray(O.gen_points_progression(curve, k0, d, n))
    zero = set(range(2, n, zero_every)) if zero_every else set()
    for i in zero:
        points[i * pb:(i + 1) * pb] = bytes(pb)
    r = M.CURVES[curve]["r"]
    rnd = random.Random(seed)
    sc = [rnd.getrandbits(256) for _ in range(n)]
    sc[:6] = [0, 1, r - 1, r, (1 << 256) - 1, (1 << 255) + 12345][: min(6, n)]
    return bytes(points), sc, (k0, d, zero, r)


def _closed(curve, sc, meta, m=None):
    k0, d, zero, r = meta
    m = len(sc) if m is None else m
    t = sum(s * (k0 + i * d) for i, s in enumerate(sc[:m]) if i not in zero) % r
    return O.g1_mul_gen(curve, t.to_bytes(32, "little"))


@pytest.mark.parametrize("curve", [1, 2])
@pytest.mark.parametrize("n,c", [(100, 0), (100, 7), (5000, 0), (5000, 9), (5000, 16), ((1 << 14) + 3, 0),
                                 ((1 << 14) + 3, 13)])
def test_table_msm_matches_plain_and_closed_form(D, curve, n, c):
    points, sc, meta = _case(curve, n, n + c, zero_every=97)
    pp = D.PreparedPoints(curve, points)
    tab = D.PointsTable(pp, window_bits=c)
    d_sc = D.to_device(b"".join(s.to_bytes(32, "little") for s in sc))
    print(f"table curve={curve} n={n} c={tab.c}", flush=True)
    got = tab.msm(d_sc).cpu().numpy().tobytes()
    plain = D.msm_prepared(pp, d_sc).cpu().numpy().tobytes()
    assert got == plain
    assert got[:O.POINT_BYTES[curve]] == _closed(curve, sc, meta)


@pytest.mark.parametrize("curve", [1, 2])
def test_table_prefixes(D, curve):
    """MSMs over the first m points of one table (the KZG SRS prefix use)."""
    n = 6000
    points, sc, meta = _case(curve, n, 3)
    pp = D.PreparedPoints(curve, points)
    tab = D.PointsTable(pp)
    d_sc = D.to_device(b"".join(s.to_bytes(32, "little") for s in sc))
    for m in (1, 17, 4095, 4096, 5999, 6000):
        got = tab.msm(d_sc, n=m).cpu().numpy().tobytes()
        assert got[:O.POINT_BYTES[curve]] == _closed(curve, sc, meta, m), m


@pytest.mark.parametrize("curve", [1, 2])
def test_table_one_bucket_for_every_digit(D, curve):
    """Every window digit of every scalar is 5: all n * rows entries land in
    one bucket, so its carry run spans thousands of chunks."""
    n, c = 5000, 12
    points, _, meta = _case(curve, n, 9)
    rows = 256 // c + (1 if 256 % c == 0 else 0)
    s = sum(5 << (c * w) for w in range(rows - 1)) % (1 << 256)
    sc = [s] * n
    pp = D.PreparedPoints(curve, points)
    tab = D.PointsTable(pp, window_bits=c)
    d_sc = D.to_device(b"".join(v.to_bytes(32, "little") for v in sc))
    got = tab.msm(d_sc).cpu().numpy().tobytes()
    assert got == D.msm_prepared(pp, d_sc).cpu().numpy().tobytes()
    assert got[:O.POINT_BYTES[curve]] == _closed(curve, sc, meta)


def test_table_argument_errors(D):
    import ctypes

    import crypto_amd

    lib = crypto_amd.lib.require_device()
    points, sc, _ = _case(2, 64, 1)
    pp = D.PreparedPoints(2, points)
    tab = D.PointsTable(pp, window_bits=8)
    d_sc = D.to_device(bytes(32 * 65))
    out = D.to_device(bytes(144))
    # more scalars than table points; bad window bits
    assert lib.kinet_hip_msm_table(2, d_sc.data_ptr(), tab.table.data_ptr(), pp.inf.data_ptr(), 64, 65, 8,
                                   out.data_ptr(), None) == -2
    assert lib.kinet_hip_msm_table(2, d_sc.data_ptr(), tab.table.data_ptr(), pp.inf.data_ptr(), 64, 64, 4,
                                   out.data_ptr(), None) == -2
    assert lib.kinet_hip_points_table_build(2, pp.pts.data_ptr(), pp.inf.data_ptr(), 64, 19,
                                            tab.table.data_ptr(), None) == -2
    assert lib.kinet_hip_table_bytes(3, 64, 8) == 0 and lib.kinet_hip_table_bytes(2, 64, 8) == 33 * 64 * 32 * 4
    assert lib.kinet_hip_msm_table(2, ctypes.c_void_p(0), tab.table.data_ptr(), None, 64, 0, 8, out.data_ptr(),
                                   None) == 0
    torch = pytest.importorskip("torch")
    torch.cuda.synchronize()
    assert out.cpu().numpy().tobytes()[:96] == bytes(96)  # empty MSM: the identity
