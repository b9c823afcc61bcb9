"""The serving scheduler (crypto_amd/csrc/serve_core.h: continuous batching,
HBM-budget admission, solo re-runs after an error flag, failure isolation)
on the CPU, through the host check's fake executor, which injects launch and
staging failures, device failures and out-of-range evaluations
(tests/hostcheck/hostcheck.cpp: FakeExec, hc_serve_scenario).  The same
scheduler runs the HIP library's serving loop (serve.hip)."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "crypto_amd", "lib", "libkinet_hostcheck.so")
STATS = ("submitted", "completed", "failed", "batches", "max_batch_requests", "peak_inflight_bytes", "budget_waits",
         "rerun_batches", "launch_failures")


@pytest.fixture(scope="module")
def hc():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "hostcheck")], check=True)
    lib = ctypes.CDLL(LIB)
    lib.hc_serve_scenario.restype = ctypes.c_int
    return lib


def run(hc, reqs, lanes=3, max_batch=128, max_elems=1 << 21, budget=8 << 30, launch_fail=(), device_fail=(),
        latency_us=200):
    """reqs: [(curve, n, bad)] -> (status, request id seen in each commitment, batch, stats)."""
    m = len(reqs)
    I = ctypes.c_int
    arr = lambda t, v: (t * max(1, len(v)))(*v)  # noqa: E731
    status, commit8, batch = (I * m)(), (ctypes.c_uint64 * m)(), (I * m)()
    st = (ctypes.c_uint64 * 9)()
    hc.hc_serve_scenario(ctypes.c_uint32(lanes), ctypes.c_uint32(max_batch), ctypes.c_uint64(max_elems),
                         ctypes.c_uint64(budget), ctypes.c_size_t(m), arr(I, [c for c, _, _ in reqs]),
                         arr(ctypes.c_uint32, [n for _, n, _ in reqs]), arr(I, [int(b) for _, _, b in reqs]),
                         arr(I, list(launch_fail) + [-1]), arr(I, list(device_fail) + [-1]), I(latency_us),
                         status, commit8, batch, st)
    return list(status), list(commit8), list(batch), dict(zip(STATS, st))


def test_mixed_stream_batches_by_curve(hc):
    reqs = [((1 if i % 3 else 2), 1 << (4 + i % 6), False) for i in range(40)]
    status, ids, batch, st = run(hc, reqs)
    assert status == [0] * 40
    assert ids == [i + 1 for i in range(40)]  # every request got its own result
    assert st["completed"] == 40 and st["failed"] == 0 and st["batches"] < 40
    for b in set(batch):  # a batch holds one curve
        assert len({reqs[i][0] for i in range(40) if batch[i] == b}) == 1


def test_bad_request_fails_alone_after_a_rerun(hc):
    reqs = [(2, 1024, i == 3) for i in range(10)]
    status, ids, batch, st = run(hc, reqs, lanes=1)
    assert status[3] == -1
    assert all(status[i] == 0 and ids[i] == i + 1 for i in range(10) if i != 3)
    assert st["rerun_batches"] == 1 and st["failed"] == 1 and st["completed"] == 9
    assert st["batches"] == 11  # the batch, then every request alone
    assert len(set(batch)) == 10


@pytest.mark.parametrize("kind", ["launch", "device"])
def test_failed_batch_fails_only_its_requests(hc, kind):
    """A batch whose staging or launch throws (kind launch: an allocation or
    launch failure), or whose device work fails, fails its own requests with
    CRYPTO_ERR_BACKEND; batches before and after it complete."""
    reqs = [(1, 2048, False) for _ in range(12)]
    inj = {"launch_fail": (1,)} if kind == "launch" else {"device_fail": (1,)}
    status, ids, batch, st = run(hc, reqs, lanes=1, max_batch=4, **inj)
    failed = [i for i in range(12) if batch[i] == 1]
    assert len(failed) == 4
    assert all(status[i] == -4 for i in failed)
    assert all(status[i] == 0 and ids[i] == i + 1 for i in range(12) if i not in failed)
    assert st["failed"] == 4 and st["completed"] == 8
    assert st["launch_failures"] == (1 if kind == "launch" else 0)


def test_admission_keeps_inflight_within_budget(hc):
    """Batches of 4 x 2048 elements (fake footprint 256 KiB each) on three
    lanes: a budget of 2.5 batches admits at most two at a time; a budget
    below one batch still runs every batch, one at a time."""
    reqs = [(1, 2048, False) for _ in range(24)]
    batch_bytes = 4 * 2048 * 32
    status, _, _, st = run(hc, reqs, lanes=3, max_batch=4, budget=int(2.5 * batch_bytes), latency_us=2000)
    assert status == [0] * 24
    assert st["peak_inflight_bytes"] <= 2.5 * batch_bytes
    assert st["budget_waits"] > 0
    status, _, _, st = run(hc, reqs, lanes=3, max_batch=4, budget=1, latency_us=2000)
    assert status == [0] * 24
    assert st["peak_inflight_bytes"] == batch_bytes
    assert st["budget_waits"] > 0


def test_batch_limits(hc):
    reqs = [(2, 2048, False) for _ in range(10)]
    _, _, batch, st = run(hc, reqs, lanes=1, max_batch=3, latency_us=1000)
    assert st["max_batch_requests"] <= 3
    _, _, batch, st = run(hc, reqs, lanes=1, max_elems=4096, latency_us=1000)
    assert st["max_batch_requests"] <= 2


def test_submit_refusals(hc):
    """Sizes that are not powers of two, above the SRS size or above the batch
    element limit, and unknown curves are refused at submission."""
    status, _, _, st = run(hc, [(1, 3, False), (3, 16, False), (2, 16, False), (1, 1 << 22, False)],
                           max_elems=1 << 21)
    assert status[0] == -1 and status[1] == -1 and status[2] == 0 and status[3] == -1
    assert st["submitted"] == 1


def test_host_build_clean_under_asan_and_ubsan():
    """The host check (field, curve, quad and GLV code, the serving scheduler
    with injected failures) built with AddressSanitizer and
    UndefinedBehaviorSanitizer runs its identities with no report."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "hostcheck"), "sanitize"], check=True)
    exe = os.path.join(ROOT, "crypto_amd", "lib", "sancheck")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out
    assert "sancheck: ok" in out
    assert "AddressSanitizer" not in out and "runtime error" not in out, out
