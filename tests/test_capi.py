"""The C-ABI library loads and exports every function include/*.h declares.
No compute calls: on a host without a GPU only the error path is exercised."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "kinet", "*", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
            if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
                names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared_functions()
    for required in ("gpukit_multi_pippenger", "gpukit_multi_pippenger_hip", "kinet_hip_msm_prepared"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import crypto_amd

    path = crypto_amd.library_path()
    if not os.path.exists(path):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_silent_cpu_route_without_device():
    """Without a usable device the ABI reports an error; nothing is computed on the CPU."""
    import crypto_amd

    if not os.path.exists(crypto_amd.library_path()):
        pytest.skip("HIP library not built")
    lib = crypto_amd.library()
    if lib.kinet_hip_available():
        pytest.skip("a HIP device is present; covered by the gpu tests")
    out = ctypes.create_string_buffer(64)
    rc = lib.gpukit_multi_pippenger(1, bytes(32), bytes(64), 1, out)
    assert rc == -3  # GPUKIT_ERR_BACKEND
    with pytest.raises(crypto_amd.HipUnavailable):
        crypto_amd.multi_pippenger(1, bytes(32), bytes(64))


def test_eip2537_msm_rejects_without_device_work():
    """bls12_381_g1_msm checks its encodings before any device call: malformed
    input returns 0 here too, and a well-formed one returns 0 without a device
    (no CPU route)."""
    import crypto_amd

    if not os.path.exists(crypto_amd.library_path()):
        pytest.skip("HIP library not built")
    lib = crypto_amd.library()
    rx, ry = ctypes.create_string_buffer(64), ctypes.create_string_buffer(64)
    assert lib.bls12_381_g1_msm(rx, ry, bytes(159), 159) == 0  # not a whole entry
    bad = bytearray(160)
    bad[0] = 1  # non-zero padding byte
    assert lib.bls12_381_g1_msm(rx, ry, bytes(bad), 160) == 0
    assert lib.bls12_381_g1_msm(rx, ry, b"", 0) == 1 and rx.raw == bytes(64) and ry.raw == bytes(64)
    if not lib.kinet_hip_available():
        assert lib.bls12_381_g1_msm(rx, ry, bytes(160), 160) == 0
