"""GPU tests of the commitment serving loop (kzg_server_*) and of the batched
kzg_commit_batch, through the C ABI: a request's commitment must be the same
bytes whether it is computed alone (kzg_commit) or inside any batch."""
import random

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

TAU = 0x1D2C3B4A59687766554433221100FFEEDDCCBBAA99887766554433221100


@pytest.fixture(scope="module")
def env():
    import crypto_amd
    from crypto_amd import serve

    crypto_amd.lib.require_device()
    srs = {1: crypto_amd.KzgSrs.generate(1, TAU, 4096), 2: crypto_amd.KzgSrs.generate(2, TAU, 4096)}
    return crypto_amd, serve, srs


def _evals(curve, seed, n):
    return np.frombuffer(O.gen_scalars(curve, seed, n), dtype=np.uint8).copy()


def test_serve_alone_vs_mix(env):
    """A mixed stream (both curves, sizes 1..4096, arbitrary interleaving)
    through the server gives each request exactly its kzg_commit bytes."""
    ca, serve, srs = env
    rnd = random.Random(7)
    reqs = []
    for i in range(60):
        curve = rnd.choice([1, 2])
        n = 1 << rnd.randint(0, 12)
        reqs.append((curve, _evals(curve, 5000 + i, n)))
    server = serve.KzgServer(srs, lanes=3, max_batch=16)
    print("serve alone vs mix: 60 requests", flush=True)
    tickets = [server.submit(c, ev) for c, ev in reqs]
    got = [server.wait(t)[0] for t in tickets]
    st = server.stats()
    server.close()
    for (curve, ev), g in zip(reqs, got):
        assert g == srs[curve].commit(ev.tobytes())
    assert st["completed"] == 60 and st["failed"] == 0
    assert st["batches"] < 60  # requests were batched
    # the smallest ones also against the oracle
    for (curve, ev), g in list(zip(reqs, got))[:6]:
        if ev.size // 32 <= 256:
            assert g == O.kzg_commit(curve, srs[curve].export(0, ev.size // 32), ev.tobytes())


def test_serve_bad_request_fails_alone(env):
    """An evaluation >= r fails only its own request; its batch mates succeed."""
    ca, serve, srs = env
    curve = 2
    good = [_evals(curve, 600 + i, 1024) for i in range(6)]
    bad = _evals(curve, 700, 1024)
    bad[32 * 5:32 * 6] = 0xFF
    server = serve.KzgServer(srs, lanes=1, max_batch=64)
    print("serve bad request", flush=True)
    tks = [server.submit(curve, ev) for ev in good[:3]] + [server.submit(curve, bad)] + \
          [server.submit(curve, ev) for ev in good[3:]]
    results = []
    for tk in tks:
        try:
            results.append(server.wait(tk)[0])
        except ca.KinetError as e:
            results.append(e.code)
    server.close()
    assert results[3] == -1
    for ev, r in zip(good, results[:3] + results[4:]):
        assert r == srs[curve].commit(ev.tobytes())


def test_serve_tiny_budget_still_completes(env):
    """With an HBM budget smaller than any batch, batches run one at a time
    (a batch is admitted when nothing is in flight) and all complete."""
    ca, serve, srs = env
    server = serve.KzgServer(srs, lanes=3, max_batch=4, hbm_budget=1)
    reqs = [(1, _evals(1, 800 + i, 2048)) for i in range(12)]
    print("serve tiny budget", flush=True)
    tks = [server.submit(c, ev) for c, ev in reqs]
    got = [server.wait(t)[0] for t in tks]
    st = server.stats()
    server.close()
    for (curve, ev), g in zip(reqs, got):
        assert g == srs[curve].commit(ev.tobytes())
    assert st["completed"] == 12
    assert st["peak_inflight_bytes"] > 0


def test_serve_poisson_short(env):
    """A short open-loop Poisson run completes every request with results equal
    to kzg_commit, and reports latencies."""
    ca, serve, srs = env
    pool = {1: _evals(1, 900, 4096), 2: _evals(2, 901, 4096)}

    def make(rng):
        curve = 2 if rng.random() < 0.7 else 1
        n = 1 << int(rng.integers(6, 13))
        off = int(rng.integers(0, 4096 - n + 1))
        return curve, pool[curve][32 * off:32 * (off + n)]

    server = serve.KzgServer(srs)
    print("serve poisson short", flush=True)
    done, wall = serve.poisson_load(server, make, rate=400.0, duration=0.5, seed=3)
    server.close()
    assert len(done) > 50
    assert all(lat > 0 for _, _, lat, _, _ in done)
    for curve, n, _, commit, ev in done[:40]:
        assert commit == srs[curve].commit(ev.tobytes())


@pytest.mark.parametrize("curve", [1, 2])
def test_commit_batch_mixed_sizes_and_handles(env, curve):
    """kzg_commit_batch groups by SRS handle and sorts by size internally; the
    output order and bytes must match the per-call kzg_commit."""
    ca, serve, srs = env
    other = ca.KzgSrs.generate(curve, TAU + 1, 1024)
    handles, evs = [], []
    rnd = random.Random(curve)
    for i in range(20):
        h = srs[curve] if i % 3 else other
        n = 1 << rnd.randint(0, 10)
        handles.append(h)
        evs.append(O.gen_scalars(curve, 1200 + i, n))
    print(f"commit batch mixed curve={curve}", flush=True)
    got = ca.kzg_commit_batch(handles, evs)
    for h, ev, g in zip(handles, evs, got):
        assert g == h.commit(ev)
    other.close()


def test_serve_and_batch_with_srs_window_tables():
    """SRSs holding several window tables (kzg_srs_precompute with different
    window bits): batched, served and single commitments keep the bytes of the
    table-free SRS, for every job size (jobs near a table's window bits use
    it; the others run ordinary windows over the tables' points)."""
    import crypto_amd
    from crypto_amd import serve

    crypto_amd.lib.require_device()
    plain = {1: crypto_amd.KzgSrs.generate(1, TAU, 8192), 2: crypto_amd.KzgSrs.generate(2, TAU, 8192)}
    tabled = {1: crypto_amd.KzgSrs.generate(1, TAU, 8192), 2: crypto_amd.KzgSrs.generate(2, TAU, 8192)}
    for curve in (1, 2):
        b0 = tabled[curve].precompute(8)
        b1 = tabled[curve].precompute(12)
        assert b1 > b0 > 0
        assert tabled[curve].precompute(12) == b1  # already there
    rnd = random.Random(11)
    reqs = []
    for i in range(40):
        curve = rnd.choice([1, 2])
        n = 1 << rnd.randint(0, 13)
        reqs.append((curve, _evals(curve, 9000 + i, n)))
    want = [plain[c].commit(ev.tobytes()) for c, ev in reqs]
    for (c, ev), w in zip(reqs, want):
        assert tabled[c].commit(ev.tobytes()) == w
    for curve in (1, 2):
        sub = [(ev, w) for (c, ev), w in zip(reqs, want) if c == curve]
        got = crypto_amd.kzg_commit_batch([tabled[curve]] * len(sub), [ev.tobytes() for ev, _ in sub])
        assert got == [w for _, w in sub]
    server = serve.KzgServer(tabled, lanes=2, max_batch=16)
    tickets = [server.submit(c, ev) for c, ev in reqs]
    got = [server.wait(t)[0] for t in tickets]
    server.close()
    assert got == want
    for s in list(plain.values()) + list(tabled.values()):
        s.close()


def test_commit_batch_graph_replays_read_new_inputs(env):
    """kzg_commit_batch with the same job sizes again and again: once the shape
    recurs (the fourth call by default) the commit pipeline (H2D, inverse
    NTTs, batched MSM, D2H) is captured and later calls replay the graph;
    every call's commitments equal kzg_commit's for that call's evaluations."""
    ca, serve, srs = env
    sizes = [16, 16, 256, 1024, 4096]
    for rep in range(8):
        evs = [_evals(2, 9000 + 10 * rep + i, n).tobytes() for i, n in enumerate(sizes)]
        print(f"commit batch replay {rep}", flush=True)
        got = ca.kzg_commit_batch([srs[2]] * len(sizes), evs)
        assert got == [srs[2].commit(e) for e in evs], rep
        if rep == 2:  # eager work on the device between replays
            srs[1].commit(_evals(1, 77, 2048).tobytes())


def test_commit_batch_graph_after_workspace_growth(env):
    """A captured commit graph is replayed only while the lane's buffers and
    MSM workspace keep their addresses: sizes A five times (captured at the
    fourth use, replayed at the fifth), then larger sizes B (which grow the
    lane's workspace), then A again -- A must be recaptured, not replayed over
    freed scratch -- and every call's commitments equal kzg_commit's."""
    ca, serve, srs = env
    plan_a = [1024, 1024, 2048]
    plan_b = [4096] * 48  # more than any batch of the earlier tests: the workspace grows
    for rep, sizes in enumerate([plan_a] * 5 + [plan_b] + [plan_a] * 2):
        evs = [_evals(2, 9500 + 10 * rep + i, n).tobytes() for i, n in enumerate(sizes)]
        print(f"commit batch after workspace growth: call {rep}, sizes {sizes}", flush=True)
        assert ca.kzg_commit_batch([srs[2]] * len(sizes), evs) == [srs[2].commit(e) for e in evs], rep
