"""Python mirror of include/kinet/crypto/kzg_serve.h (the commitment serving
loop runs natively in the HIP library; this module only submits, waits and
drives open-loop Poisson load for measurements)."""
from __future__ import annotations

import ctypes
import threading
import time

import numpy as np

from .lib import CRYPTO_STATUS, KinetError, compressed_bytes, require_device


class ServerConfig(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("max_batch_elems", ctypes.c_uint64), ("hbm_budget", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "failed", "batches", "max_batch_requests",
                                               "peak_inflight_bytes", "budget_waits", "rerun_batches",
                                               "launch_failures")]


class KzgServer:
    """kzg_server_*: srs maps curve id (1 BN254, 2 BLS12-381) to a KzgSrs."""

    def __init__(self, srs: dict, lanes: int = 0, max_batch: int = 0, max_batch_elems: int = 0,
                 hbm_budget: int = 0):
        lib = require_device()
        self._srs = srs  # keep the handles alive
        cfg = ServerConfig(lanes, max_batch, max_batch_elems, hbm_budget)
        h = ctypes.c_void_p()
        rc = lib.kzg_server_create(srs[1].handle if 1 in srs else None, srs[2].handle if 2 in srs else None,
                                   ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise KinetError("kzg_server_create", rc, CRYPTO_STATUS)
        self.handle = h.value
        self._keep = {}

    def submit(self, curve: int, evals) -> int:
        """evals: bytes or a C-contiguous numpy uint8 array (kept alive until wait)."""
        arr = evals if isinstance(evals, np.ndarray) else np.frombuffer(evals, dtype=np.uint8)
        tk = require_device().kzg_server_submit(self.handle, curve, arr.ctypes.data, arr.size // 32)
        if tk < 0:
            raise KinetError("kzg_server_submit", int(tk), CRYPTO_STATUS)
        self._keep[tk] = (curve, arr)
        return tk

    def wait(self, ticket: int):
        """Returns (commitment bytes, submit ns, done ns)."""
        curve, _ = self._keep[ticket]
        out = ctypes.create_string_buffer(compressed_bytes(curve))
        t0, t1 = ctypes.c_uint64(), ctypes.c_uint64()
        rc = require_device().kzg_server_wait(self.handle, ticket, out, ctypes.byref(t0), ctypes.byref(t1))
        del self._keep[ticket]
        if rc != 0:
            raise KinetError("kzg_server_wait", rc, CRYPTO_STATUS)
        return out.raw, t0.value, t1.value

    def stats(self) -> dict:
        st = ServerStats()
        require_device().kzg_server_get_stats(self.handle, ctypes.byref(st))
        return {k: int(getattr(st, k)) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            require_device().kzg_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_load(server: KzgServer, make_request, rate: float, duration: float, seed: int = 1):
    """Open-loop load: arrivals with exponential gaps (mean 1/rate) for
    `duration` seconds; make_request(rng) -> (curve, evals array).  A waiter
    thread collects completions.  Returns ([(curve, n, latency_s, commitment,
    evals)], wall seconds), latency measured by the server from submit to
    completion, in submission order, so every result can be checked against
    its request."""
    rng = np.random.default_rng(seed)
    tickets = []
    done = []
    cond = threading.Condition()
    finished = [False]

    def waiter():
        k = 0
        while True:
            with cond:
                while k >= len(tickets) and not finished[0]:
                    cond.wait()
                if k >= len(tickets):
                    return
                tk, curve, ev = tickets[k]
            commit, t0, t1 = server.wait(tk)
            done.append((curve, ev.size // 32, (t1 - t0) * 1e-9, commit, ev))
            k += 1

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    start = time.perf_counter()
    next_t = start
    while True:
        next_t += rng.exponential(1.0 / rate)
        if next_t - start > duration:
            break
        now = time.perf_counter()
        if next_t > now:
            time.sleep(next_t - now)
        curve, ev = make_request(rng)
        tk = server.submit(curve, ev)
        with cond:
            tickets.append((tk, curve, ev))
            cond.notify()
    with cond:
        finished[0] = True
        cond.notify()
    th.join()
    return done, time.perf_counter() - start
