"""ctypes mirror of the HIP library's C ABI.

Names, argument meaning and return codes follow the reference's C entry points
(gpukit/include/kinet/gpukit/multi_pippenger.h:57-92, c-abi/crypto.h:60-66);
Python callers get bytes in, bytes out, and an exception for a non-OK status.
"""
from __future__ import annotations

import ctypes
import os

try:
    # torch ships its own libamdhip64.so.7 (same SONAME as /opt/rocm's).  Loading
    # torch first makes the HIP library bind to that one runtime, so device
    # pointers and streams from torch and from this library share a context.
    import torch  # noqa: F401
except ImportError:  # the C ABI itself needs no torch
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libkinet_crypto_hip.so")

CURVE_BN254_G1 = 1
CURVE_BLS12_381_G1 = 2
CURVE_BANDERWAGON = 3
FIELD_BLS12_381_FR = 0
FIELD_BN254_FR = 1

GPUKIT_OK = 0
GPUKIT_STATUS = {0: "GPUKIT_OK", -1: "GPUKIT_ERR_NULL_ARG", -2: "GPUKIT_ERR_BAD_SIZE", -3: "GPUKIT_ERR_BACKEND",
                 -4: "GPUKIT_ERR_DISPATCH", -5: "GPUKIT_ERR_NOTIMPL", -6: "GPUKIT_ERR_DETERMINISM"}
CRYPTO_STATUS = {0: "CRYPTO_OK", -1: "CRYPTO_ERR_INPUT", -2: "CRYPTO_ERR_LENGTH", -3: "CRYPTO_ERR_VERIFY",
                 -4: "CRYPTO_ERR_BACKEND", -5: "CRYPTO_ERR_NOTIMPL", -6: "CRYPTO_ERR_INTERNAL"}


class KinetNttGroup(ctypes.Structure):
    """kinet_ntt_group of include/kinet/crypto/hip_runtime.h."""
    _fields_ = [("log_n", ctypes.c_uint32), ("count", ctypes.c_uint32), ("offset", ctypes.c_uint64)]


class HipUnavailable(RuntimeError):
    """The HIP library is missing or no HIP device is usable."""


class KinetError(RuntimeError):
    def __init__(self, fn, code, table):
        self.code = code
        super().__init__(f"{fn} returned {code} ({table.get(code, '?')}): {last_error()}")


_lib = None


def library_path() -> str:
    return _LIB_PATH


def library():
    """Loads the HIP library (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise HipUnavailable(f"HIP library not built: {_LIB_PATH} (run __graft_entry__.build())")
        lib = ctypes.CDLL(_LIB_PATH)
        vp, sz, u32, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int
        lib.kinet_hip_last_error.restype = ctypes.c_char_p
        lib.kinet_hip_available.restype = i32
        lib.gpukit_multi_pippenger_hip.argtypes = [u32, ctypes.c_char_p, ctypes.c_char_p, sz, ctypes.c_char_p]
        lib.gpukit_multi_pippenger_hip.restype = i32
        lib.bls12_381_g1_msm.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz]
        lib.bls12_381_g1_msm.restype = i32
        lib.kinet_hip_points_prepare.argtypes = [u32, vp, sz, vp, vp, vp]
        lib.kinet_hip_points_bytes.argtypes = [u32, sz]
        lib.kinet_hip_points_bytes.restype = sz
        lib.kinet_hip_msm_prepared.argtypes = [u32, vp, vp, vp, sz, i32, i32, vp, vp]
        lib.kinet_hip_msm_prepared_g1.argtypes = [u32, vp, vp, vp, sz, i32, i32, vp, vp]
        lib.kinet_hip_msm_prepared_noalloc.argtypes = [u32, vp, vp, vp, sz, vp, vp]
        lib.kinet_hip_points_check_g1.argtypes = [u32, vp, vp, sz, vp, vp]
        lib.kinet_hip_points_decompress.argtypes = [u32, vp, sz, vp, vp, vp]
        lib.kinet_hip_g1_fixed_base_mul.argtypes = [u32, vp, sz, vp, vp]
        lib.kinet_hip_valu_peak.argtypes = [i32, ctypes.POINTER(ctypes.c_double)]
        for name in ("gpukit_batch_inv_secp256k1_fp_hip", "gpukit_batch_inv_bn254_fp_hip",
                     "gpukit_batch_inv_bls12_381_fp_hip"):
            getattr(lib, name).argtypes = [ctypes.c_char_p, ctypes.c_char_p, sz]
            getattr(lib, name).restype = i32
        lib.kinet_hip_batch_inv_fp.argtypes = [u32, vp, vp, sz, vp, vp]
        lib.kinet_hip_batch_inv_fp.restype = i32
        for name in ("ntt_fr_forward", "ntt_fr_inverse"):
            getattr(lib, name).argtypes = [u32, ctypes.c_void_p, sz]
            getattr(lib, name).restype = i32
        lib.kinet_hip_ntt_fr.argtypes = [u32, i32, vp, vp, vp, sz, vp, vp]
        pp = ctypes.POINTER(ctypes.c_void_p)
        lib.kzg_srs_create.argtypes = [u32, ctypes.c_char_p, sz, pp]
        lib.kzg_open.argtypes = [vp, ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
        lib.kzg_open.restype = i32
        lib.kzg_open_coeffs.argtypes = [vp, ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
        lib.kzg_open_coeffs.restype = i32
        lib.kzg_commit_coeffs.argtypes = [vp, ctypes.c_char_p, sz, ctypes.c_char_p]
        lib.kzg_commit_coeffs.restype = i32
        lib.kzg_srs_generate.argtypes = [u32, ctypes.c_char_p, sz, pp]
        lib.kzg_srs_destroy.argtypes = [vp]
        lib.kzg_srs_destroy.restype = None
        lib.kzg_srs_size.argtypes = [vp]
        lib.kzg_srs_size.restype = sz
        lib.kzg_srs_curve.argtypes = [vp]
        lib.kzg_srs_curve.restype = u32
        lib.kzg_srs_export.argtypes = [vp, sz, sz, ctypes.c_char_p]
        lib.kzg_srs_precompute.argtypes = [vp, i32]
        lib.kzg_srs_table_bytes.argtypes = [vp]
        lib.kzg_srs_table_bytes.restype = sz
        lib.kzg_commit.argtypes = [vp, ctypes.c_char_p, sz, ctypes.c_char_p]
        lib.kzg_commit_batch.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(sz), sz,
                                         ctypes.c_char_p]
        lib.kzg_msm.argtypes = [vp, ctypes.c_char_p, sz, ctypes.c_char_p]
        lib.kzg_ctx_create.argtypes = [vp, pp]
        lib.kzg_ctx_destroy.argtypes = [vp]
        lib.kzg_ctx_destroy.restype = None
        lib.kzg_commit_async.argtypes = [vp, vp, vp, sz, i32, vp, vp]
        lib.kinet_hip_profile_enable.argtypes = [i32]
        lib.kinet_hip_profile_enable.restype = None
        lib.kinet_hip_msm_set_reduce_segment.argtypes = [i32]
        lib.kinet_hip_msm_set_reduce_segment.restype = i32
        lib.kinet_hip_profile_collect.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32), i32]
        lib.kinet_hip_ntt_fr_batch.argtypes = [u32, i32, vp, vp, vp, sz, sz, vp, vp]
        lib.kinet_hip_msm_batch_create.argtypes = [u32, ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_uint64),
                                                   ctypes.POINTER(u32), sz, i32, pp]
        lib.kinet_hip_msm_batch_destroy.argtypes = [vp]
        lib.kinet_hip_msm_batch_destroy.restype = None
        lib.kinet_hip_msm_batch_window_bits.argtypes = [vp, sz]
        lib.kinet_hip_msm_batch_run.argtypes = [vp, vp, vp, vp, vp, i32, vp]
        lib.kinet_hip_job_set_create.argtypes = [u32, ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_uint64),
                                                 ctypes.POINTER(u32), sz, u32, i32, ctypes.POINTER(KinetNttGroup), sz,
                                                 pp]
        lib.kinet_hip_job_set_run.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        lib.kinet_hip_job_set_destroy.argtypes = [vp]
        lib.kinet_hip_job_set_destroy.restype = None
        lib.kinet_hip_job_set_window_bits.argtypes = [vp, sz]
        lib.kinet_hip_msm_set_window_table.argtypes = [u32, i32, ctypes.POINTER(i32), i32]
        lib.kinet_hip_msm_window_bits.argtypes = [u32, sz, i32]
        lib.kinet_hip_table_window_bits.argtypes = [sz]
        lib.kinet_hip_table_bytes.argtypes = [u32, sz, i32]
        lib.kinet_hip_table_bytes.restype = sz
        lib.kinet_hip_points_table_build.argtypes = [u32, vp, vp, sz, i32, vp, vp]
        lib.kinet_hip_msm_table.argtypes = [u32, vp, vp, vp, sz, sz, i32, vp, vp]
        ip = ctypes.POINTER(i32)
        lib.kinet_hip_tune_msm.argtypes = [u32, i32, i32, i32]
        lib.kinet_hip_tune_ntt.argtypes = [u32, i32, i32]
        lib.kinet_hip_tune_save.argtypes = [ctypes.c_char_p]
        lib.kinet_hip_tune_load.argtypes = [ctypes.c_char_p]
        lib.kinet_hip_tune_clear.argtypes = []
        lib.kinet_hip_tune_clear.restype = None
        lib.kinet_hip_tune_get_msm.argtypes = [u32, i32, i32, ip, ip]
        lib.kinet_hip_tune_get_ntt.argtypes = [u32, i32, ip, ip, ip]
        lib.kzg_server_create.argtypes = [vp, vp, vp, pp]
        lib.kzg_server_destroy.argtypes = [vp]
        lib.kzg_server_destroy.restype = None
        lib.kzg_server_submit.argtypes = [vp, u32, vp, sz]
        lib.kzg_server_submit.restype = ctypes.c_int64
        lib.kzg_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_uint64)]
        lib.kzg_server_get_stats.argtypes = [vp, vp]
        _lib = lib
    return _lib


def last_error() -> str:
    if _lib is None:
        return ""
    msg = _lib.kinet_hip_last_error()
    return msg.decode() if msg else ""


def require_device():
    lib = library()
    if not lib.kinet_hip_available():
        raise HipUnavailable(f"no usable HIP device: {last_error()}")
    return lib


def point_bytes(curve: int) -> int:
    return 96 if curve == CURVE_BLS12_381_G1 else 64


def compressed_bytes(curve: int) -> int:
    return 48 if curve == CURVE_BLS12_381_G1 else 32


def ntt_fr(field: int, elems: bytes, inverse: bool = False) -> bytes:
    """ntt_fr_forward / ntt_fr_inverse on 32-byte LE canonical elements."""
    lib = require_device()
    n = len(elems) // 32
    buf = ctypes.create_string_buffer(bytes(elems), len(elems))
    fn = lib.ntt_fr_inverse if inverse else lib.ntt_fr_forward
    rc = fn(field, ctypes.cast(buf, ctypes.c_void_p), n)
    if rc != 0:
        raise KinetError(fn.__name__, rc, CRYPTO_STATUS)
    return buf.raw[: len(elems)]


class KzgSrs:
    """Device-resident SRS (kzg_srs_create / kzg_srs_generate)."""

    def __init__(self, handle, curve: int, n: int):
        self.handle, self.curve, self.n = handle, curve, n

    @classmethod
    def from_points(cls, curve: int, points: bytes) -> "KzgSrs":
        lib = require_device()
        n = len(points) // point_bytes(curve)
        h = ctypes.c_void_p()
        rc = lib.kzg_srs_create(curve, points, n, ctypes.byref(h))
        if rc != 0:
            raise KinetError("kzg_srs_create", rc, CRYPTO_STATUS)
        return cls(h, curve, n)

    @classmethod
    def generate(cls, curve: int, tau: int, n: int) -> "KzgSrs":
        lib = require_device()
        h = ctypes.c_void_p()
        rc = lib.kzg_srs_generate(curve, tau.to_bytes(32, "little"), n, ctypes.byref(h))
        if rc != 0:
            raise KinetError("kzg_srs_generate", rc, CRYPTO_STATUS)
        return cls(h, curve, n)

    def export(self, first: int = 0, count: int | None = None) -> bytes:
        count = self.n - first if count is None else count
        out = ctypes.create_string_buffer(point_bytes(self.curve) * count)
        rc = library().kzg_srs_export(self.handle, first, count, out)
        if rc != 0:
            raise KinetError("kzg_srs_export", rc, CRYPTO_STATUS)
        return out.raw

    def commit(self, evals: bytes) -> bytes:
        out = ctypes.create_string_buffer(compressed_bytes(self.curve))
        rc = library().kzg_commit(self.handle, evals, len(evals) // 32, out)
        if rc != 0:
            raise KinetError("kzg_commit", rc, CRYPTO_STATUS)
        return out.raw

    def open(self, evals: bytes, z: int) -> tuple[bytes, int]:
        """Opening proof at z (kzg_open): (compressed proof, y = f(z))."""
        proof = ctypes.create_string_buffer(compressed_bytes(self.curve))
        y = ctypes.create_string_buffer(32)
        rc = library().kzg_open(self.handle, evals, len(evals) // 32, z.to_bytes(32, "little"), proof, y)
        if rc != 0:
            raise KinetError("kzg_open", rc, CRYPTO_STATUS)
        return proof.raw, int.from_bytes(y.raw, "little")

    def commit_coeffs(self, coeffs: bytes) -> bytes:
        """Commitment from coefficients c_i of f = sum c_i X^i (kzg_commit_coeffs)."""
        out = ctypes.create_string_buffer(compressed_bytes(self.curve))
        rc = library().kzg_commit_coeffs(self.handle, coeffs, len(coeffs) // 32, out)
        if rc != 0:
            raise KinetError("kzg_commit_coeffs", rc, CRYPTO_STATUS)
        return out.raw

    def open_coeffs(self, coeffs: bytes, z: int) -> tuple[bytes, int]:
        """Opening proof at any z from coefficients (kzg_open_coeffs): (compressed proof, y = f(z))."""
        proof = ctypes.create_string_buffer(compressed_bytes(self.curve))
        y = ctypes.create_string_buffer(32)
        rc = library().kzg_open_coeffs(self.handle, coeffs, len(coeffs) // 32, z.to_bytes(32, "little"), proof, y)
        if rc != 0:
            raise KinetError("kzg_open_coeffs", rc, CRYPTO_STATUS)
        return proof.raw, int.from_bytes(y.raw, "little")

    def msm(self, scalars: bytes) -> bytes:
        out = ctypes.create_string_buffer(point_bytes(self.curve))
        rc = library().kzg_msm(self.handle, scalars, len(scalars) // 32, out)
        if rc != 0:
            raise KinetError("kzg_msm", rc, CRYPTO_STATUS)
        return out.raw

    def precompute(self, window_bits: int = 0) -> int:
        """Builds the SRS's window table (kzg_srs_precompute); returns its device bytes."""
        rc = library().kzg_srs_precompute(self.handle, window_bits)
        if rc != 0:
            raise KinetError("kzg_srs_precompute", rc, CRYPTO_STATUS)
        return library().kzg_srs_table_bytes(self.handle)

    def close(self):
        if self.handle:
            library().kzg_srs_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kzg_commit_batch(srs_list, evals_list) -> list[bytes]:
    lib = require_device()
    count = len(srs_list)
    arr_srs = (ctypes.c_void_p * count)(*[s.handle for s in srs_list])
    arr_ev = (ctypes.c_char_p * count)(*evals_list)
    arr_n = (ctypes.c_size_t * count)(*[len(e) // 32 for e in evals_list])
    cb = compressed_bytes(srs_list[0].curve) if count else 0
    out = ctypes.create_string_buffer(max(1, cb * count))
    rc = lib.kzg_commit_batch(arr_srs, arr_ev, arr_n, count, out)
    if rc != 0:
        raise KinetError("kzg_commit_batch", rc, CRYPTO_STATUS)
    return [out.raw[cb * i: cb * (i + 1)] for i in range(count)]


def bls12_381_g1_msm(xycs: bytes) -> tuple[bool, bytes, bytes]:
    """EIP-2537 G1 MSM (include/kinet/crypto/bls.h; the reference's
    bls12_381_g1_msm, bls/c-abi/c_bls_pairing.h:153-156): xycs is N tuples of
    x[64] || y[64] || c[32] (big-endian); returns (ok, rx[64], ry[64]) with the
    reference's convention of ok = False on any encoding, curve or subgroup
    failure."""
    lib = require_device()
    rx, ry = ctypes.create_string_buffer(64), ctypes.create_string_buffer(64)
    ok = lib.bls12_381_g1_msm(rx, ry, bytes(xycs), len(xycs)) == 1
    return ok, rx.raw, ry.raw


BATCH_INV_BYTES = {0: 32, 1: 32, 2: 48}  # base-field element bytes: secp256k1, BN254, BLS12-381
_BATCH_INV = {0: "gpukit_batch_inv_secp256k1_fp_hip", 1: "gpukit_batch_inv_bn254_fp_hip",
              2: "gpukit_batch_inv_bls12_381_fp_hip"}


def batch_inverse(curve: int, elems: bytes) -> bytes:
    """Inverses of base-field elements (LE records) through
    gpukit_batch_inv_<field>_fp_hip; KinetError (GPUKIT_ERR_BAD_SIZE) when an
    element is zero."""
    lib = require_device()
    nb = BATCH_INV_BYTES[curve]
    if len(elems) % nb:
        raise ValueError(f"elements are {nb}-byte records")
    out = ctypes.create_string_buffer(max(1, len(elems)))
    fn = getattr(lib, _BATCH_INV[curve])
    rc = fn(elems, out, len(elems) // nb)
    if rc != GPUKIT_OK:
        raise KinetError(fn.__name__, rc, GPUKIT_STATUS)
    return out.raw[: len(elems)]


def multi_pippenger(curve: int, scalars: bytes, points: bytes) -> bytes:
    """sum s_i P_i through gpukit_multi_pippenger_hip (the reference's wire
    format: 32-byte LE scalars, BE affine points, all-zero = identity; for
    Banderwagon the result's identity is (0, 1) and bad inputs raise with
    GPUKIT_ERR_BAD_SIZE)."""
    lib = require_device()
    n = len(scalars) // 32
    if len(scalars) != 32 * n or len(points) != point_bytes(curve) * n:
        raise ValueError("scalars must be n*32 bytes and points n*point_bytes(curve) bytes")
    out = ctypes.create_string_buffer(point_bytes(curve))
    fn = lib.gpukit_multi_pippenger_hip
    rc = fn(curve, scalars, points, n, out)
    if rc != GPUKIT_OK:
        raise KinetError(fn.__name__, rc, GPUKIT_STATUS)
    return out.raw
