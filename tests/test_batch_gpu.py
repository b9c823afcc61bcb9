"""GPU parity of the batched paths (kinet_hip_msm_batch_*, kinet_hip_ntt_fr_batch)
through the C ABI: every job of a batch must give the bytes the same MSM or
transform gives alone, and the closed form / oracle value.  Bar: bit-exact."""
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    import crypto_amd
    from crypto_amd import device

    crypto_amd.lib.require_device()
    return device


def _progression(curve, k0, d, n):
    return O.gen_points_progression(curve, k0, d, n)


def _closed(curve, scalars, ks):
    """(sum s_i k_i mod r) G for scalars (bytes) against points k_i G."""
    r = M.CURVES[curve]["r"]
    t = 0
    for i, k in enumerate(ks):
        t += int.from_bytes(scalars[32 * i:32 * i + 32], "little") * k
    return O.g1_mul_gen(curve, (t % r).to_bytes(32, "little"))


def _scalar_block(rnd, n, wide):
    """n scalars; with wide=True the first few are >= 2^255 (top carry)."""
    sc = [rnd.getrandbits(256 if wide else 250) for _ in range(n)]
    if wide:
        sc[:min(n, 5)] = [(1 << 256) - 1 - i for i in range(min(n, 5))]
    return b"".join(s.to_bytes(32, "little") for s in sc)


@pytest.mark.parametrize("curve", [1, 2])
@pytest.mark.parametrize("window_bits", [0, 5, 8, 13, 16])
def test_msm_batch_matches_closed_form(D, curve, window_bits):
    k0, d = 0xC0FFEE, 0x1357913579
    npts = 5000
    points = bytearray(_progression(curve, k0, d, npts))
    ks = [(k0 + i * d) for i in range(npts)]
    pb = O.POINT_BYTES[curve]
    for i in (7, 99, 1500, 2500, 4999):  # identity points (all-zero wire bytes)
        points[i * pb:(i + 1) * pb] = bytes(pb)
        ks[i] = 0
    points = bytes(points)
    sizes = [1, 0, 2, 3, 64, 100, 777, 1024, 3000, 5000, 17]
    rnd = random.Random(curve * 100 + window_bits)
    blocks = [_scalar_block(rnd, n, wide=(j % 2 == 0)) for j, n in enumerate(sizes)]
    scalars = b"".join(blocks)
    pp = D.PreparedPoints(curve, points)
    b = D.MsmBatch(curve, sizes, window_bits=window_bits)
    print(f"msm batch curve={curve} c={window_bits} jobs={len(sizes)}", flush=True)
    out = b.run(pp, D.to_device(scalars)).cpu().numpy().tobytes()
    rec = pb + pb // 2
    for j, n in enumerate(sizes):
        want = _closed(curve, blocks[j], ks[:n]) if n else bytes(pb)
        got = out[j * rec:(j + 1) * rec]
        assert got[:pb] == want, (j, n)
        assert got[pb:] == O.compress(curve, want), (j, n)
        if window_bits:
            assert b.window_bits(j) == window_bits
    b.close()


@pytest.mark.parametrize("curve", [1, 2])
def test_msm_batch_equals_single_msms(D, curve):
    """Each job of a batch gives exactly the record of the same MSM alone, with
    explicit point and scalar offsets (jobs over different point ranges,
    scalars stored in reverse job order)."""
    npts = 4096
    points = _progression(curve, 31337, 0xABCDEF, npts)
    pp = D.PreparedPoints(curve, points)
    rnd = random.Random(7 + curve)
    sizes = [4096, 1000, 1, 2048, 333, 64]
    pt_off = [0, 96, 4095, 1024, 3000, 7]
    total = sum(sizes)
    scalars = _scalar_block(rnd, total, wide=True)
    sc_off, acc = [], total
    for n in sizes:  # reverse order
        acc -= n
        sc_off.append(acc)
    d_sc = D.to_device(scalars)
    b = D.MsmBatch(curve, sizes, sc_off=sc_off, pt_off=pt_off)
    print(f"msm batch offsets curve={curve}", flush=True)
    out = b.run(pp, d_sc).cpu().numpy().tobytes()
    rec = O.POINT_BYTES[curve] * 3 // 2
    import torch

    for j, n in enumerate(sizes):
        # the same MSM alone: points pt_off.., scalars sc_off..
        pts_j = D.PreparedPoints(curve, points[pt_off[j] * O.POINT_BYTES[curve]:(pt_off[j] + n) * O.POINT_BYTES[curve]])
        sc_j = d_sc[sc_off[j] * 32:(sc_off[j] + n) * 32].clone()
        alone = D.msm_prepared(pts_j, sc_j, n=n, window_bits=b.window_bits(j)).cpu().numpy().tobytes()
        assert out[j * rec:(j + 1) * rec] == alone, j
        want = O.msm(curve, scalars[sc_off[j] * 32:(sc_off[j] + n) * 32],
                     points[pt_off[j] * O.POINT_BYTES[curve]:(pt_off[j] + n) * O.POINT_BYTES[curve]])
        assert alone[:O.POINT_BYTES[curve]] == want, j
    del torch
    b.close()


def test_msm_batch_graph_replays(D):
    """Graph mode: every replay equals the eager run on the same inputs, with
    eager work and synchronisations between replays, and a replay after the
    scalars were overwritten in place follows the new scalars."""
    import torch

    curve = 1
    npts = 2048
    k0, d = 99, 0x5555
    points = _progression(curve, k0, d, npts)
    ks = [k0 + i * d for i in range(npts)]
    pp = D.PreparedPoints(curve, points)
    sizes = [2048, 1500, 700, 1024, 5]
    rnd = random.Random(5)
    sc1 = _scalar_block(rnd, sum(sizes), wide=False)
    sc2 = _scalar_block(rnd, sum(sizes), wide=True)
    d_sc = D.to_device(sc1)
    b = D.MsmBatch(curve, sizes)
    s = torch.cuda.Stream()
    out_e = torch.empty(len(sizes) * b.record, dtype=torch.uint8, device="cuda")
    out_g = torch.empty_like(out_e)
    with torch.cuda.stream(s):
        b.run(pp, d_sc, out=out_e, use_graph=False, stream=s)
        eager = out_e.cpu().numpy().tobytes()
        for rep in range(3):
            print(f"graph replay {rep}", flush=True)
            out_g.fill_(0xAA)
            s.synchronize()
            b.run(pp, d_sc, out=out_g, use_graph=True, stream=s)
            assert out_g.cpu().numpy().tobytes() == eager, rep
            D.msm_prepared(pp, d_sc, n=100, stream=s)  # eager work on the same stream
            torch.cuda.synchronize()
        d_sc.copy_(D.to_device(sc2))
        s.synchronize()
        b.run(pp, d_sc, out=out_g, use_graph=True, stream=s)
        replay = out_g.cpu().numpy().tobytes()
        b.run(pp, d_sc, out=out_e, use_graph=False, stream=s)
        assert replay == out_e.cpu().numpy().tobytes()
    off = 0
    pb = O.POINT_BYTES[curve]
    for j, n in enumerate(sizes):
        assert replay[j * b.record:j * b.record + pb] == _closed(curve, sc2[32 * off:32 * (off + n)], ks[:n]), j
        off += n
    b.close()


def test_msm_batch_log_uniform_mix(D):
    """A small copy of the batched workload: 48 BN254 MSMs with log-uniform
    sizes in [2^10, 2^14] over one shared point prefix."""
    curve = 1
    rnd = random.Random(2024)
    sizes = [int(2 ** rnd.uniform(10, 14)) for _ in range(48)]
    npts = max(sizes)
    k0, d = 0xDEADBEEF, 0x1234567
    points = _progression(curve, k0, d, npts)
    ks = [k0 + i * d for i in range(npts)]
    pp = D.PreparedPoints(curve, points)
    scalars = O.gen_scalars(curve, 77, sum(sizes))
    b = D.MsmBatch(curve, sizes)
    print(f"log-uniform batch jobs={len(sizes)} points={sum(sizes)}", flush=True)
    out = b.run(pp, D.to_device(scalars)).cpu().numpy().tobytes()
    off = 0
    pb = O.POINT_BYTES[curve]
    for j, n in enumerate(sizes):
        assert out[j * b.record:j * b.record + pb] == _closed(curve, scalars[32 * off:32 * (off + n)], ks[:n]), j
        off += n
    b.close()


@pytest.mark.parametrize("field", [0, 1])
@pytest.mark.parametrize("logn,count", [(0, 3), (3, 5), (11, 4), (16, 2), (21, 3)])
def test_ntt_batch_matches_single(D, field, logn, count):
    """count transforms in one batch equal the transforms run one by one; the
    small ones also equal the oracle."""
    import torch

    n = 1 << logn
    curve = 2 if field == 0 else 1
    data = O.gen_scalars(curve, 1000 + logn, n * count)
    d_in = D.to_device(data)
    print(f"ntt batch field={field} logn={logn} count={count}", flush=True)
    for inverse in (False, True):
        got = D.ntt_batch(field, d_in, n, count, inverse=inverse).cpu().numpy().tobytes()
        for t in range(count):
            one = D.ntt_batch(field, d_in[t * n * 32:(t + 1) * n * 32].clone(), n, 1,
                              inverse=inverse).cpu().numpy().tobytes()
            assert got[t * n * 32:(t + 1) * n * 32] == one, (t, inverse)
            if logn <= 11:
                assert one == O.ntt(field, data[t * n * 32:(t + 1) * n * 32], inverse=inverse), (t, inverse)
    del torch


def test_ntt_batch_rejects_noncanonical(D):
    """An element >= r anywhere in the batch raises the error flag."""
    import torch

    n, count = 1024, 3
    data = bytearray(O.gen_scalars(2, 5, n * count))
    data[(2 * n + 17) * 32:(2 * n + 18) * 32] = b"\xff" * 32
    from crypto_amd.lib import KinetError

    with pytest.raises(KinetError):
        D.ntt_batch(0, D.to_device(bytes(data)), n, count)
    del torch


@pytest.mark.parametrize("curve,field", [(1, 1), (2, 0)])
def test_job_set_one_graph_matches_per_job(D, curve, field):
    """kinet_hip_job_set_*: uneven MSMs and NTT groups in one graph launch give
    each job's bytes from the per-job entries (kinet_hip_msm_prepared,
    kinet_hip_ntt_fr), on the first run (capture) and on replays with eager
    work and new inputs in between; *d_err is cleared by every run and set by
    a non-canonical NTT input."""
    import torch

    rnd = random.Random(curve * 7 + field)
    sizes = [1, 3, 100, 4097, 20000, 777]
    npts = max(sizes)
    k0, d = 1234567, 0x1000001
    pp = D.PreparedPoints(curve, _progression(curve, k0, d, npts))
    logs = [(4, 3), (10, 2), (13, 1)]  # (log2 n, count)
    off, groups = 0, []
    for lg, cnt in logs:
        groups.append((lg, cnt, off))
        off += cnt << lg
    d_tmp = torch.empty(off * 32, dtype=torch.uint8, device="cuda")
    d_nout = torch.empty_like(d_tmp)
    d_err = torch.zeros(1, dtype=torch.int32, device="cuda")
    out = torch.empty(len(sizes) * (D.point_bytes(curve) * 3 // 2), dtype=torch.uint8, device="cuda")
    rec = D.point_bytes(curve) * 3 // 2
    stream = torch.cuda.Stream()
    jset = D.JobSet(curve, sizes, field, groups)
    try:
        for rep in range(3):
            sc = b"".join(rnd.getrandbits(254).to_bytes(32, "little") for _ in range(sum(sizes)))
            vals = b"".join(rnd.randrange(M.BLS_R if field == 0 else M.BN_R).to_bytes(32, "little")
                            for _ in range(off))
            d_sc = D.to_device(sc)
            d_nin = D.to_device(vals)
            if rep == 0:
                keep_sc, keep_nin = d_sc, d_nin
            else:  # the same buffers, refilled: replays read them afresh
                keep_sc.copy_(d_sc)
                keep_nin.copy_(d_nin)
            torch.cuda.synchronize()
            print(f"job set curve={curve} field={field} run={rep}", flush=True)
            jset.run(pp, keep_sc, out, keep_nin, d_nout, d_tmp, d_err, stream=stream)
            assert int(d_err.item()) == 0
            host = out.cpu().numpy().tobytes()
            o = 0
            for j, n in enumerate(sizes):
                one = D.msm_prepared(pp, keep_sc[32 * o:32 * (o + n)], n=n).cpu().numpy().tobytes()
                assert host[rec * j:rec * (j + 1)] == one, (rep, j)
                o += n
            for lg, cnt, go in groups:
                for t in range(cnt):
                    a = 32 * (go + (t << lg))
                    b = a + (32 << lg)
                    want = O.ntt(field, vals[a:b])
                    assert d_nout[a:b].cpu().numpy().tobytes() == want, (rep, lg, t)
        bad = bytearray(vals)
        bad[32 * 5:32 * 6] = bytes([0xFF]) * 32
        keep_nin.copy_(D.to_device(bytes(bad)))
        torch.cuda.synchronize()
        jset.run(pp, keep_sc, out, keep_nin, d_nout, d_tmp, d_err, stream=stream)
        assert int(d_err.item()) != 0
    finally:
        jset.close()
