"""GPU parity at the configured full sizes, through the C ABI.

Each result is checked against a closed form that shares no code with the
device path (O(n) scalar-field arithmetic in the C oracle, then one scalar
multiplication of the generator):
  - MSM over points k_i G:      sum s_i P_i = (sum s_i k_i mod r) G
  - KZG commit over [tau^i] G:  C = [f(tau)] G, f(tau) from the evaluations by
                                the barycentric formula (no NTT)
  - NTT:                        X_k = sum_j x_j w^(jk) by Horner at w^k
Points k_i G come from the library's fixed-base kernel (pinned against the
oracle's k G in test_msm_gpu.py::test_fixed_base_generator).
"""
import random

import numpy as np
import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu

R_BW = 0x1CFB69D4CA675F520CCE760202687600FF8F87007419047174FD06B52876E7E1  # Banderwagon group order


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.lib.require_device()
    return crypto_amd


def _fixed_base_points(ca, curve, ks: bytes):
    import torch
    from crypto_amd import device as D

    n = len(ks) // 32
    out = torch.empty(n * ca.point_bytes(curve), dtype=torch.uint8, device="cuda")
    rc = ca.library().kinet_hip_g1_fixed_base_mul(curve, D.to_device(ks).data_ptr(), n, out.data_ptr(),
                                                  D.stream_handle())
    assert rc == 0, rc
    torch.cuda.synchronize()
    return out


def _check_point(curve, rec: bytes, want: bytes):
    pb = O.POINT_BYTES[curve]
    assert rec[:pb] == want
    assert rec[pb:pb + O.COMP_BYTES[curve]] == O.compress(curve, want)
    # the compressed form decodes to the closed-form affine point (x and y in
    # the 64/96-byte big-endian wire format of the reference)
    assert O.decompress(curve, rec[pb:pb + O.COMP_BYTES[curve]]) == want


@pytest.mark.parametrize("curve,g1", [(2, True), (2, False), (1, True), (1, False)])
def test_msm_2p20_closed_form(ca, curve, g1):
    """The headline MSM size, N = 2^20, on the GLV (G1) path and the plain
    path: seeded uniform scalars below r, plus edge scalars 0, 1, r - 1, r,
    2^255 and 2^256 - 1 in the first slots."""
    from crypto_amd import device as D
    from crypto_amd import inputs

    n = 1 << 20
    r = M.CURVES[curve]["r"]
    ks = inputs.scalars_bytes(curve, 1001 + curve, n)
    sc = bytearray(inputs.scalars_bytes(curve, 2002 + curve, n))
    for i, v in enumerate([0, 1, r - 1, r, 1 << 255, (1 << 256) - 1]):
        sc[32 * i:32 * i + 32] = v.to_bytes(32, "little")
    sc = bytes(sc)
    print(f"full-size MSM curve={curve} n=2^20 g1={g1}", flush=True)
    pp = D.PreparedPoints(curve, _fixed_base_points(ca, curve, ks))
    out = D.msm_prepared(pp, D.to_device(sc), g1=g1).cpu().numpy().tobytes()
    t = O.fr_dot(O.fr_of_curve(curve), sc, ks)
    _check_point(curve, out, O.g1_mul_gen(curve, t.to_bytes(32, "little")))


@pytest.mark.parametrize("curve,g1,host", [(2, True, False), (1, False, True), (3, False, False)])
def test_msm_beyond_one_pipeline(ca, curve, g1, host):
    """n = 2^25 + 3 points: two slices (the second of 3 points) summed on the
    device, through the prepared entries (BLS12-381 on the GLV path,
    Banderwagon) and the host-buffer entry gpukit_multi_pippenger_hip
    (BN254); identity points and edge scalars at both ends of the first
    slice."""
    from crypto_amd import device as D
    from crypto_amd import inputs

    n = (1 << 25) + 3
    cut = 1 << 25
    r = M.CURVES[curve]["r"] if curve != 3 else R_BW
    ks = bytearray(inputs.scalars_bytes(curve, 3003 + curve, n))
    sc = bytearray(inputs.scalars_bytes(curve, 4004 + curve, n))
    for i in (0, cut - 1, cut):  # identity points (k = 0)
        ks[32 * i:32 * i + 32] = bytes(32)
    for i, v in ((1, r - 1), (cut - 2, 1), (cut + 1, 0), (n - 1, r - 1)):
        sc[32 * i:32 * i + 32] = v.to_bytes(32, "little")
    ks, sc = bytes(ks), bytes(sc)
    print(f"sliced MSM curve={curve} n=2^25+3 g1={g1} host={host}", flush=True)
    wire = _fixed_base_points(ca, curve, ks)
    if host:
        out = ca.multi_pippenger(curve, sc, wire.cpu().numpy().tobytes())
    else:
        pp = D.PreparedPoints(curve, wire)
        out = D.msm_prepared(pp, D.to_device(sc), g1=g1).cpu().numpy().tobytes()
    del wire
    if curve == 3:
        want = O.bw_mul_gen(O.fr_dot(2, sc, ks))
        assert out[:64] == want
        if not host:
            assert out[64:96] == O.bw_compress(want)
        return
    t = O.fr_dot(O.fr_of_curve(curve), sc, ks)
    want = O.g1_mul_gen(curve, t.to_bytes(32, "little"))
    if host:
        assert out == want
    else:
        _check_point(curve, out, want)


@pytest.mark.parametrize("curve", [1, 2])
def test_kzg_commit_2p22_closed_form(ca, curve):
    """Full KZG commit at 2^22 evaluations: iNTT on the device, MSM against a
    2^22-point SRS [tau^i] G generated on the device, compressed output,
    against [f(tau)] G."""
    from crypto_amd import inputs

    n = 1 << 22
    r = M.CURVES[curve]["r"]
    tau = random.Random(22 + curve).randrange(2, r)
    evals = inputs.scalars_bytes(curve, 3003 + curve, n)
    print(f"full-size KZG commit curve={curve} n=2^22", flush=True)
    srs = ca.KzgSrs.generate(curve, tau, n)
    try:
        got = srs.commit(evals)
    finally:
        srs.close()
    ft = O.fr_lagrange(O.fr_of_curve(curve), evals, tau)
    want = O.g1_mul_gen(curve, ft.to_bytes(32, "little"))
    assert got == O.compress(curve, want)
    assert O.decompress(curve, got) == want


def test_batched_set_every_job_closed_form(ca):
    """The full batched job set -- 256 BN254 MSMs of 2^10 .. 2^18 points over a
    shared 2^18-point SRS prefix and 256 BN254-Fr NTTs of 2^12 .. 2^20
    elements -- as ONE HIP graph launch (kinet_hip_job_set_run: the MSM batch
    pipeline and the NTT groups on a forked branch), replayed with eager work
    between replays.  Every MSM job against its closed form; every NTT job at
    one random output against Horner, plus the round trip iNTT(NTT(x)) = x;
    every replay equal to the first run."""
    import torch
    from crypto_amd import device as D
    from crypto_amd import inputs
    from crypto_amd import workloads as W

    curve, field = 1, 1
    msm_n, ntt_log = W.batched_set_sizes()
    ks = inputs.scalars_bytes(curve, 4242, W.BATCH_SRS_POINTS)
    sc = inputs.scalars_bytes(curve, 4243, sum(msm_n))
    print(f"batched set: {len(msm_n)} MSMs ({sum(msm_n)} points), {len(ntt_log)} NTTs, one graph", flush=True)
    pp = D.PreparedPoints(curve, _fixed_base_points(ca, curve, ks))
    d_sc = D.to_device(sc)
    off_of, total, groups = W.ntt_layout(ntt_log)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(4244)
    d_in = torch.randint(0, 256, (total * 32,), dtype=torch.uint8, device="cuda", generator=gen)
    d_in.view(-1, 32)[:, 31] &= 0x1F  # < 2^253 < r
    d_out = torch.empty_like(d_in)
    d_tmp = torch.empty_like(d_in)
    d_err = torch.ones(1, dtype=torch.int32, device="cuda")  # the run clears it
    out = torch.empty(len(msm_n) * 96, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.Stream()
    jset = D.JobSet(curve, msm_n, field, [(lg, len(jobs), off_of[jobs[0]]) for lg, jobs in groups.items()])
    try:
        jset.run(pp, d_sc, out, d_in, d_out, d_tmp, d_err, stream=stream)
        first_msm, first_ntt = out.clone(), d_out.clone()
        assert int(d_err.item()) == 0
        for rep in range(2):
            out.zero_()
            d_out.zero_()
            torch.cuda.synchronize()
            jset.run(pp, d_sc, out, d_in, d_out, d_tmp, d_err, stream=stream)
            assert torch.equal(out, first_msm) and torch.equal(d_out, first_ntt), rep
    finally:
        jset.close()
    # the MSM batch alone (kinet_hip_msm_batch_*) gives the same records
    batch = D.MsmBatch(curve, msm_n)
    try:
        assert torch.equal(batch.run(pp, d_sc), first_msm)
    finally:
        batch.close()
    host = first_msm.cpu().numpy().tobytes()
    off = 0
    for j, n in enumerate(msm_n):
        t = O.fr_dot(field, sc[32 * off:32 * (off + n)], ks[:32 * n])
        _check_point(curve, host[96 * j:96 * (j + 1)], O.g1_mul_gen(curve, t.to_bytes(32, "little")))
        off += n

    d_back = torch.empty_like(d_in)
    for lg, jobs in groups.items():
        o = 32 * off_of[jobs[0]]
        cnt = len(jobs) * (32 << lg)
        D.ntt_batch(field, d_out[o:o + cnt], 1 << lg, len(jobs), inverse=True, d_out=d_back[o:o + cnt],
                    d_tmp=d_tmp[o:o + cnt], d_err=d_err, sync=False)
    torch.cuda.synchronize()
    assert int(d_err.item()) == 0
    assert torch.equal(d_back, d_in)
    hin, hout = d_in.cpu().numpy(), d_out.cpu().numpy()
    rnd = random.Random(4245)
    r = M.BN_R
    for j, lg in enumerate(ntt_log):
        n, o = 1 << lg, off_of[j]
        k = rnd.randrange(n)
        x = pow(O.root_of_unity(field, n), k, r)
        got = int.from_bytes(hout[32 * (o + k):32 * (o + k + 1)].tobytes(), "little")
        assert O.fr_eval(field, hin[32 * o:32 * (o + n)].tobytes(), x) == got, (j, n, k)


def test_serving_mix_full_sizes_every_result_closed_form(ca):
    """The serving mix at its configured sizes -- 70% BLS12-381 n = 4096, 30%
    BN254 n = 2^14 .. 2^20, each request a different window of evaluations --
    under open-loop Poisson arrivals; every result against its closed form."""
    from crypto_amd import inputs, serve
    from crypto_amd import workloads as W

    tau = 0x0BADC0FFEE0DDF00D5EED5EED5EED5EED5EED5EED5EED5EED5EED5EED
    srs = {2: ca.KzgSrs.generate(2, tau, W.SERVE_BLS_N), 1: ca.KzgSrs.generate(1, tau, 1 << W.SERVE_BN254_LOG[1])}
    pool = {2: inputs.scalars(2, 501, W.SERVE_BLS_N + 512).view(np.uint8).reshape(-1).copy(),
            1: inputs.scalars(1, 502, (1 << 20) + 4096).view(np.uint8).reshape(-1).copy()}
    server = serve.KzgServer(srs)
    print("serving mix at full sizes: Poisson 150/s for 2 s", flush=True)
    try:
        done, _ = serve.poisson_load(server, lambda g: W.serving_request(g, pool, distinct=True), rate=150.0,
                                     duration=2.0, seed=503)
        st = server.stats()
    finally:
        server.close()
        for s in srs.values():
            s.close()
    sizes = {(c, n) for c, n, _, _, _ in done}
    assert (1, 1 << 14) in sizes and (1, 1 << 20) in sizes and (2, 4096) in sizes
    assert st["failed"] == 0 and st["completed"] == len(done)
    for curve, n, lat, commit, ev in done:
        ft = O.fr_lagrange(O.fr_of_curve(curve), ev.tobytes(), tau)
        assert commit == O.compress(curve, O.g1_mul_gen(curve, ft.to_bytes(32, "little"))), (curve, n)
