"""kinet_hip_msm_prepared_noalloc (scratch in a static device array, no
allocation) against kinet_hip_msm_prepared (library-allocated workspace),
through the C ABI: byte-equal records on two calls in a row, the second with
a smaller input, on every curve; the size limit; the empty MSM; calls on two
streams."""
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

MAX_N = 1 << 16


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.lib.require_device()
    return crypto_amd


def _points(ca, curve, n, seed):
    import torch
    from crypto_amd import device as D
    from crypto_amd import inputs

    ks = inputs.scalars_bytes(curve, seed, n)
    wire = torch.empty(n * ca.point_bytes(curve), dtype=torch.uint8, device="cuda")
    assert ca.library().kinet_hip_g1_fixed_base_mul(curve, D.to_device(ks).data_ptr(), n, wire.data_ptr(),
                                                    D.stream_handle()) == 0
    torch.cuda.synchronize()
    return D.PreparedPoints(curve, wire), ks


def _noalloc(ca, pp, d_sc, n, stream=None, sync=True):
    import torch
    from crypto_amd import device as D

    pb = ca.point_bytes(pp.curve)
    out = torch.full((pb + pb // 2,), 0xEE, dtype=torch.uint8, device="cuda")
    rc = ca.library().kinet_hip_msm_prepared_noalloc(pp.curve, d_sc.data_ptr(), pp.pts.data_ptr(),
                                                     pp.inf.data_ptr(), n, out.data_ptr(), D.stream_handle(stream))
    if sync:
        torch.cuda.synchronize()
    return rc, out


@pytest.mark.parametrize("curve", [0, 1, 2])
def test_noalloc_matches_allocating_entry_twice(ca, curve):
    from crypto_amd import device as D
    from crypto_amd import inputs

    pp, ks = _points(ca, curve, MAX_N, 7000 + curve)
    for rep, n in enumerate((MAX_N, 3000)):
        sc = inputs.scalars_bytes(curve, 7100 + 10 * curve + rep, n)
        d_sc = D.to_device(sc)
        print(f"noalloc curve={curve} call={rep} n={n}", flush=True)
        rc, got = _noalloc(ca, pp, d_sc, n)
        assert rc == 0, ca.lib.last_error()
        want = D.msm_prepared(pp, d_sc, n=n).cpu().numpy().tobytes()
        assert got.cpu().numpy().tobytes() == want
    if curve != 0:  # and the last one against the closed form
        t = O.fr_dot(O.fr_of_curve(curve), sc, ks[:32 * n])
        assert want[:ca.point_bytes(curve)] == O.g1_mul_gen(curve, t.to_bytes(32, "little"))


def test_noalloc_limits_and_empty(ca):
    from crypto_amd import device as D
    from crypto_amd import inputs

    curve = 2
    pp, _ = _points(ca, curve, 64, 7200)
    d_sc = D.to_device(inputs.scalars_bytes(curve, 7201, 64))
    print("noalloc limits", flush=True)
    rc, out = _noalloc(ca, pp, d_sc, MAX_N + 1)
    assert rc == -2  # GPUKIT_ERR_BAD_SIZE, before any device work
    assert bool((out == 0xEE).all())
    assert "exceeds" in ca.lib.last_error()
    rc, out = _noalloc(ca, pp, d_sc, 0)
    assert rc == 0
    want = D.msm_prepared(pp, d_sc, n=0).cpu().numpy().tobytes()
    assert out.cpu().numpy().tobytes() == want
    assert want[:96] == bytes(96) and want[96] == 0xC0
    assert ca.library().kinet_hip_msm_prepared_noalloc(3, d_sc.data_ptr(), pp.pts.data_ptr(), pp.inf.data_ptr(), 1,
                                                       out.data_ptr(), D.stream_handle()) == -3


def test_noalloc_two_streams(ca):
    """Back-to-back calls on two streams share the array; the second waits."""
    import torch
    from crypto_amd import device as D
    from crypto_amd import inputs

    curve = 1
    pp, ks = _points(ca, curve, MAX_N, 7300)
    sc1 = inputs.scalars_bytes(curve, 7301, MAX_N)
    sc2 = inputs.scalars_bytes(curve, 7302, 5000)
    d1, d2 = D.to_device(sc1), D.to_device(sc2)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    print("noalloc two streams", flush=True)
    rc1, o1 = _noalloc(ca, pp, d1, MAX_N, stream=sa, sync=False)
    rc2, o2 = _noalloc(ca, pp, d2, 5000, stream=sb, sync=False)
    torch.cuda.synchronize()
    assert rc1 == 0 and rc2 == 0
    pb = ca.point_bytes(curve)
    for sc, o, n in ((sc1, o1, MAX_N), (sc2, o2, 5000)):
        t = O.fr_dot(1, sc, ks[:32 * n])
        assert o.cpu().numpy().tobytes()[:pb] == O.g1_mul_gen(curve, t.to_bytes(32, "little"))
