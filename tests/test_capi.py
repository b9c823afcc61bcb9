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
    for required in ("gpukit_multi_pippenger_hip", "kinet_hip_msm_prepared", "bls12_381_g1_msm", "ntt_fr_forward",
                     "kzg_commit", "kzg_commit_batch", "kinet_bn254_hip_g1_add", "kinet_hip_msm_prepared_noalloc"):
        assert required in names
    # the multi-backend dispatcher stays the reference's: this library must not
    # define a second gpukit_multi_pippenger
    assert "gpukit_multi_pippenger" not in names


def test_library_exports_every_declared_symbol():
    import crypto_amd

    path = crypto_amd.library_path()
    if not os.path.exists(path):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert not hasattr(lib, "gpukit_multi_pippenger")


def test_no_silent_cpu_route_without_device():
    """Without a usable device the ABI reports an error; nothing is computed on the CPU."""
    import crypto_amd

    if not os.path.exists(crypto_amd.library_path()):
        pytest.skip("HIP library not built")
    lib = crypto_amd.library()
    if lib.kinet_hip_available():
        pytest.skip("a HIP device is present; covered by the gpu tests")
    out = ctypes.create_string_buffer(64)
    rc = lib.gpukit_multi_pippenger_hip(1, bytes(32), bytes(64), 1, out)
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


def test_hip_backend_argument_checks_without_device():
    """gpukit_multi_pippenger_hip refuses bad arguments before touching the
    device, with the reference's status codes: NULL result or inputs, an
    unknown curve and n >= 2^32 (32-bit point indices); n = 0 is the identity
    (all-zero bytes, and (0, 1) for Banderwagon, as the reference writes it)."""
    import crypto_amd

    if not os.path.exists(crypto_amd.library_path()):
        pytest.skip("HIP library not built")
    lib = crypto_amd.library()
    out = ctypes.create_string_buffer(96)
    one = bytes(96)
    assert lib.gpukit_multi_pippenger_hip(2, one, one, 1, None) == -1
    assert lib.gpukit_multi_pippenger_hip(2, None, one, 1, out) == -1
    assert lib.gpukit_multi_pippenger_hip(4, one, one, 1, out) == -3
    assert lib.gpukit_multi_pippenger_hip(7, one, one, 1, out) == -3
    assert lib.gpukit_multi_pippenger_hip(2, one, one, 1 << 32, out) == -2
    out = ctypes.create_string_buffer(b"\xee" * 96, 96)
    assert lib.gpukit_multi_pippenger_hip(2, None, None, 0, out) == 0 and out.raw == bytes(96)
    out = ctypes.create_string_buffer(b"\xee" * 64, 64)
    assert lib.gpukit_multi_pippenger_hip(3, None, None, 0, out) == 0
    assert out.raw == bytes(63) + b"\x01"
    assert lib.gpukit_multi_pippenger_hip(3, one, one, 1 << 32, out) == -2


def test_c_caller_compiles_links_and_runs(tmp_path):
    """A C99 program including every public header compiles with -Wall
    -Werror, links against the library and gets the documented status codes
    for the calls that need no device."""
    import shutil
    import subprocess

    import crypto_amd

    path = crypto_amd.library_path()
    if not os.path.exists(path):
        pytest.skip("HIP library not built")
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "caller")
    libdir = os.path.dirname(path)
    subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_caller", "caller.c"), "-L", libdir, "-lkinet_crypto_hip",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "C caller: ok" in res.stdout


def test_blob_entries_and_setup_parser_without_device(tmp_path):
    """The EIP-4844 entries refuse malformed input before any device work:
    non-canonical blob elements and z (CRYPTO_ERR_INPUT), a missing setup
    (CRYPTO_ERR_BACKEND), and every malformed setup file (CRYPTO_ERR_INPUT)."""
    import crypto_amd

    if not os.path.exists(crypto_amd.library_path()):
        pytest.skip("HIP library not built")
    lib = crypto_amd.library()
    lib.kzg_load_trusted_setup_file.argtypes = [ctypes.c_char_p]
    r = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
    out, y = ctypes.create_string_buffer(48), ctypes.create_string_buffer(32)
    bad = bytes(32 * 5) + r.to_bytes(32, "big") + bytes(131072 - 32 * 6)
    assert lib.kzg_blob_to_commit(bad, out) == -1
    assert lib.kzg_commit_to_proof(bytes(131072), r.to_bytes(32, "big"), out, y) == -1
    assert lib.kzg_blob_to_commit(None, out) == -1
    if not lib.kinet_hip_available():
        os.environ.pop("KINET_KZG_TRUSTED_SETUP", None)
        assert lib.kzg_blob_to_commit(bytes(131072), out) == -4
    g1 = "a" * 96
    g2 = "b" * 192
    cases = {
        "empty": "",
        "count": "4095\n65\n" + (g1 + "\n") * 4095 + (g2 + "\n") * 65,
        "g2count": "4096\n1\n" + (g1 + "\n") * 4096 + g2 + "\n",
        "short": "4096\n65\n" + (g1 + "\n") * 4000,
        "hex": "4096\n65\n" + (g1 + "\n") * 4095 + "zz" * 48 + "\n" + (g2 + "\n") * 65,
        "g2flag": "4096\n2\n" + (g1 + "\n") * 4096 + ("3" + "b" * 191 + "\n") * 2,
        "trailing": "4096\n2\n" + (g1 + "\n") * 4096 + (g2 + "\n") * 2 + (g1 + "\n") * 4096 + "00\n",
        "monomial": "4096\n2\n" + (g1 + "\n") * 4096 + (g2 + "\n") * 2 + (g1 + "\n") * 10,
    }
    for name, text in cases.items():
        p = tmp_path / (name + ".txt")
        p.write_text(text)
        assert lib.kzg_load_trusted_setup_file(str(p).encode()) == -1, name
    assert lib.kzg_load_trusted_setup_file(str(tmp_path / "missing.txt").encode()) == -1
