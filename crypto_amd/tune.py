"""Schedule tables of the HIP library and its autotuner.

Schedules are chosen per size class k = floor(log2 n): for MSMs per curve and
mode (single, batched job, GLV single over G1 points; secp256k1 single and GLV) the window bits c and
the first-level bucket-reduction segment (a GLV cell of window bits 1: the
plain MSM beat the split there and runs instead); for the Fr NTT per field the split
of log2 n into one to three LDS passes.  A schedule changes speed only: the
tuner refuses any candidate whose output differs from the built-in schedule's
by a byte.  Tables persist as a text file (``save`` / ``load``); the library
also loads the file named by ``KINET_HIP_TUNE_FILE`` when it starts.

No counterpart in the reference, whose CPU Pippenger picks its window from n
alone (gpukit/cpp/cpu_reference/multi_pippenger.cpp).
"""

from __future__ import annotations

import ctypes
import os

from .lib import KinetError, library, require_device

_STATUS = {-1: "null argument", -2: "bad size", -3: "backend", -6: "schedules disagree"}

MODE_SINGLE, MODE_BATCHED, MODE_GLV = 0, 1, 2
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables", "mi355x.tune")


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise KinetError(what, rc, _STATUS)


def tune_msm(curve: int, mode: int, log_min: int, log_max: int) -> None:
    """Times the MSM candidates of size classes 2^log_min..2^log_max on the device and installs the fastest."""
    _check(require_device().kinet_hip_tune_msm(curve, mode, log_min, log_max), "tune_msm")


def tune_ntt(field: int, log_min: int, log_max: int) -> None:
    """Times every NTT pass split of size classes 2^log_min..2^log_max and installs the fastest."""
    _check(require_device().kinet_hip_tune_ntt(field, log_min, log_max), "tune_ntt")


def save(path: str) -> None:
    _check(library().kinet_hip_tune_save(os.fsencode(path)), "tune save")


def load(path: str = DEFAULT_TABLE) -> None:
    _check(library().kinet_hip_tune_load(os.fsencode(path)), "tune load")


def clear() -> None:
    library().kinet_hip_tune_clear()


def msm_cell(curve: int, mode: int, log_n: int) -> tuple[int, int]:
    """(window bits, segment) installed for the size class; 0 = built-in rule."""
    c, seg = ctypes.c_int32(), ctypes.c_int32()
    _check(library().kinet_hip_tune_get_msm(curve, mode, log_n, ctypes.byref(c), ctypes.byref(seg)), "tune get")
    return c.value, seg.value


def ntt_split(field: int, log_n: int) -> tuple[int, int, int]:
    """log2 sub-transform lengths of NTT passes 1..3 in effect for the size class."""
    p = [ctypes.c_int32() for _ in range(3)]
    _check(library().kinet_hip_tune_get_ntt(field, log_n, *(ctypes.byref(x) for x in p)), "tune get")
    return tuple(x.value for x in p)


def window_bits(curve: int, n: int, mode: int = MODE_SINGLE) -> int:
    """Window bits an n-point MSM of ``mode`` uses with the tables as installed now."""
    return library().kinet_hip_msm_window_bits(curve, n, mode)


def tune_all(log_min: int = 10, log_max: int = 22, ntt_min: int = 4, ntt_max: int = 22, verbose: bool = True) -> None:
    """Tunes every MSM mode of both curves and both NTT fields over the given size classes."""
    import time

    jobs = []
    for curve in (2, 1):
        jobs.append((f"msm curve={curve} single", lambda c=curve: tune_msm(c, MODE_SINGLE, log_min, log_max)))
        jobs.append((f"msm curve={curve} batched",
                     lambda c=curve: tune_msm(c, MODE_BATCHED, max(1, log_min - 2), min(log_max, 20))))
    for curve in (2, 1):
        jobs.append((f"msm curve={curve} glv", lambda c=curve: tune_msm(c, MODE_GLV, max(11, log_min), log_max)))
    jobs.append(("msm curve=0 single", lambda: tune_msm(0, MODE_SINGLE, log_min, log_max)))  # secp256k1
    jobs.append(("msm curve=0 glv", lambda: tune_msm(0, MODE_GLV, max(11, log_min), log_max)))
    for field in (0, 1):
        jobs.append((f"ntt field={field}", lambda f=field: tune_ntt(f, ntt_min, ntt_max)))
    for name, fn in jobs:
        t0 = time.time()
        fn()
        if verbose:
            print(f"tuned {name} in {time.time() - t0:.1f} s", flush=True)


def _main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description="Tune the MSM / NTT schedules of this device and save the table.")
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--log-min", type=int, default=10)
    ap.add_argument("--log-max", type=int, default=22)
    ap.add_argument("--ntt-min", type=int, default=4)
    ap.add_argument("--ntt-max", type=int, default=22)
    a = ap.parse_args(argv)
    clear()
    tune_all(a.log_min, a.log_max, a.ntt_min, a.ntt_max)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    save(a.out)
    print(open(a.out).read(), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
