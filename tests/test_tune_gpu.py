"""GPU tests of the autotuner: it installs a schedule for every size class it
tunes, and the MSM and NTT results under tuned (and hand-set) schedules stay
bit-exact against the oracle and the closed form."""
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu


@pytest.fixture()
def T():
    import crypto_amd
    import crypto_amd.tune as T

    crypto_amd.lib.require_device()
    T.clear()
    yield T
    T.clear()


def _progression_case(curve, n, seed):
    """Points (k0 + i d) G, random scalars, and the closed-form result."""
    k0, d = 0x1234567 + seed, 0xABCDEF1234567 + 3 * seed
    points = O.gen_points_progression(curve, k0, d, n)
    rnd = random.Random(seed)
    sc = [rnd.getrandbits(256) for _ in range(n)]
    r = M.CURVES[curve]["r"]
    t = sum(s * (k0 + i * d) for i, s in enumerate(sc)) % r
    return b"".join(s.to_bytes(32, "little") for s in sc), points, O.g1_mul_gen(curve, t.to_bytes(32, "little"))


@pytest.mark.parametrize("curve,mode,lo,hi", [(2, 0, 9, 11), (1, 0, 9, 11), (2, 2, 11, 12), (1, 2, 11, 12),
                                              (2, 1, 8, 9), (1, 1, 8, 9)])
def test_tune_msm_cells_and_results(T, curve, mode, lo, hi):
    from crypto_amd import device as D

    print(f"tune msm curve={curve} mode={mode} 2^{lo}..2^{hi}", flush=True)
    T.tune_msm(curve, mode, lo, hi)
    for lg in range(lo, hi + 1):
        c, seg = T.msm_cell(curve, mode, lg)
        if mode == T.MODE_GLV and c == 1:  # the plain MSM won this size class
            assert seg == 0 and T.window_bits(curve, 1 << lg, mode) == T.window_bits(curve, 1 << lg, T.MODE_SINGLE)
            continue
        assert 5 <= c <= 18 and c == T.window_bits(curve, 1 << lg, mode)
        assert seg in ((0,) if mode == T.MODE_BATCHED else (0, 2, 4, 8, 16))
    assert T.msm_cell(curve, mode, hi + 1) == (0, 0)
    pb = O.POINT_BYTES[curve]
    for lg in range(lo, hi + 1):
        n = (1 << lg) + 37  # same size class, not a power of two
        scalars, points, want = _progression_case(curve, n, lg)
        pp = D.PreparedPoints(curve, points)
        d_sc = D.to_device(scalars)
        if mode == T.MODE_BATCHED:
            b = D.MsmBatch(curve, [n, n], sc_off=[0, 0])
            assert b.window_bits(0) == T.msm_cell(curve, mode, lg)[0]
            out = b.run(pp, d_sc).cpu().numpy().tobytes()
            rec = b.record
            assert out[:pb] == want and out[rec:rec + pb] == want
            b.close()
        else:
            got = D.msm_prepared(pp, d_sc, g1=(mode == T.MODE_GLV)).cpu().numpy().tobytes()
            assert got[:pb] == want


def test_tune_msm_rejects_bad_ranges(T):
    import crypto_amd

    for args in ((2, 2, 9, 12), (2, 0, 0, 10), (2, 0, 12, 11), (2, 0, 20, 25), (3, 0, 10, 11),
                 (2, 3, 10, 11)):
        with pytest.raises(crypto_amd.KinetError) as e:
            T.tune_msm(*args)
        assert e.value.code == -2


@pytest.mark.parametrize("field", [0, 1])
def test_tune_ntt_then_oracle(T, field):
    import crypto_amd

    print(f"tune ntt field={field}", flush=True)
    T.tune_ntt(field, 2, 13)
    for lg in range(2, 14):
        p = T.ntt_split(field, lg)
        assert sum(p) == lg and all(0 <= v <= 11 for v in p)
    rnd = random.Random(field)
    fr = M.BLS_R if field == 0 else M.BN_R
    for lg in (2, 7, 11, 13):
        x = b"".join(rnd.randrange(fr).to_bytes(32, "little") for _ in range(1 << lg))
        assert crypto_amd.ntt_fr(field, x) == O.ntt(field, x)
        assert crypto_amd.ntt_fr(field, x, inverse=True) == O.ntt(field, x, inverse=True)


@pytest.mark.parametrize("field", [0, 1])
def test_every_ntt_split_matches_oracle(T, field, tmp_path):
    """Each pass split of log2 n = 12 (two and three passes, each pass up to
    2^11) gives the oracle's transform, forward and inverse."""
    import crypto_amd

    lg = 12
    fr = M.BLS_R if field == 0 else M.BN_R
    rnd = random.Random(7 + field)
    x = b"".join(rnd.randrange(fr).to_bytes(32, "little") for _ in range(1 << lg))
    fwd, inv = O.ntt(field, x), O.ntt(field, x, inverse=True)
    splits = [(p1, lg - p1, 0) for p1 in range(1, 12)] + [(a, b, lg - a - b) for a in range(1, 12)
                                                         for b in range(1, 12) if 1 <= lg - a - b <= 11]
    path = tmp_path / "split.tune"
    for sp in splits:
        path.write_text(f"ntt {field} {lg} {sp[0]} {sp[1]} {sp[2]}\n")
        T.load(str(path))
        assert T.ntt_split(field, lg) == sp
        assert crypto_amd.ntt_fr(field, x) == fwd, sp
        assert crypto_amd.ntt_fr(field, x, inverse=True) == inv, sp


def test_tuned_table_roundtrip(T, tmp_path):
    T.tune_msm(1, 0, 10, 10)
    T.tune_ntt(1, 10, 10)
    cell, split = T.msm_cell(1, 0, 10), T.ntt_split(1, 10)
    path = str(tmp_path / "t.tune")
    T.save(path)
    T.clear()
    assert T.msm_cell(1, 0, 10) == (0, 0)
    T.load(path)
    assert T.msm_cell(1, 0, 10) == cell and T.ntt_split(1, 10) == split
