"""ctypes bindings for the CPU oracle (oracle/liborc.so) and, when built, the
reference's own CPU code (oracle/_ref/libref.so).

Test infrastructure: only tests, the smoke check and the bench's CPU-baseline
leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIBORC = os.path.join(ORACLE_DIR, "liborc.so")
LIBREF = os.path.join(ORACLE_DIR, "_ref", "libref.so")
LIBBNREF = os.path.join(ORACLE_DIR, "_ref", "libbnref.so")

CURVE_BN254 = 1
CURVE_BLS12_381 = 2
FIELD_BLS_FP, FIELD_BLS_FR, FIELD_BN_FP, FIELD_BN_FR = 0, 1, 2, 3

CURVE_BANDERWAGON = 3
FIELD_BW_FR = 4  # Banderwagon scalar field

POINT_BYTES = {0: 64, CURVE_BN254: 64, CURVE_BLS12_381: 96, CURVE_BANDERWAGON: 64}  # 0 = secp256k1
COMP_BYTES = {CURVE_BN254: 32, CURVE_BLS12_381: 48}

_u8p = ctypes.c_char_p
_sz = ctypes.c_size_t

_orc = None
_ref = None


def _build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liborc.so"], check=True)


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(LIBORC):
            _build_oracle()
        lib = ctypes.CDLL(LIBORC)
        lib.orc_splitmix64.restype = ctypes.c_uint64
        lib.orc_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        lib.orc_gen_scalars.argtypes = [ctypes.c_int, ctypes.c_uint64, _sz, _u8p]
        lib.orc_msm.argtypes = [ctypes.c_int, _u8p, _u8p, _sz, _u8p]
        lib.orc_msm_naive.argtypes = [ctypes.c_int, _u8p, _u8p, _sz, _u8p]
        lib.orc_msm_mt.argtypes = [ctypes.c_int, _u8p, _u8p, _sz, _u8p, ctypes.c_int]
        lib.orc_ntt.argtypes = [ctypes.c_int, _u8p, _sz, ctypes.c_int]
        lib.orc_dft_naive.argtypes = [ctypes.c_int, _u8p, _u8p, _sz, ctypes.c_int]
        lib.orc_root_of_unity.argtypes = [ctypes.c_int, _sz, _u8p]
        lib.orc_kzg_commit.argtypes = [ctypes.c_int, _u8p, _u8p, _sz, _u8p]
        lib.orc_kzg_commit_closed_form.argtypes = [ctypes.c_int, _u8p, _u8p, _sz, _u8p]
        lib.orc_powers.argtypes = [ctypes.c_int, _u8p, _sz, _u8p]
        lib.orc_gen_points.argtypes = [ctypes.c_int, _u8p, _sz, _u8p]
        lib.orc_gen_points_progression.argtypes = [ctypes.c_int, _u8p, _u8p, _sz, _u8p]
        lib.orc_field_mont_consts.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.POINTER(ctypes.c_uint64)]
        lib.orc_fr_dot.argtypes = [ctypes.c_int, _u8p, _u8p, _sz, _u8p]
        lib.orc_fr_eval.argtypes = [ctypes.c_int, _u8p, _sz, _u8p, _u8p]
        lib.orc_fr_lagrange.argtypes = [ctypes.c_int, _u8p, _sz, _u8p, _u8p]
        lib.orc_fr_divide_linear.argtypes = [ctypes.c_int, _u8p, _sz, _u8p, _u8p, _u8p]
        for name, args in (("orc_bw_generator", [_u8p]), ("orc_bw_check", [_u8p]),
                           ("orc_bw_mul", [_u8p, _u8p, _u8p]), ("orc_bw_add", [_u8p, _u8p, _u8p]),
                           ("orc_bw_dbl", [_u8p, _u8p]), ("orc_bw_compress", [_u8p, _u8p]),
                           ("orc_bw_msm", [_u8p, _u8p, _sz, _u8p]), ("orc_bw_msm_naive", [_u8p, _u8p, _sz, _u8p]),
                           ("orc_bw_gen_points", [_u8p, _sz, _u8p])):
            getattr(lib, name).argtypes = args
        _orc = lib
    return _orc


_bnref = None


def bnref_available() -> bool:
    return os.path.exists(LIBBNREF)


def bnref():
    """The reference's BN254 CUDA driver built without CUDA: its CPU path
    behind kinet_bn254_cuda_{g1_add, g1_mul, fp_mul} (same layout as the HIP
    kinet_bn254_hip_* entries)."""
    global _bnref
    if _bnref is None:
        lib = ctypes.CDLL(LIBBNREF)
        for name in ("kinet_bn254_cuda_g1_add", "kinet_bn254_cuda_g1_mul", "kinet_bn254_cuda_fp_mul"):
            getattr(lib, name).argtypes = [_u8p, _u8p, _u8p, ctypes.c_uint]
            getattr(lib, name).restype = ctypes.c_int
        _bnref = lib
    return _bnref


# BN254 batch layout of the reference driver (bn254/gpu/cuda/bn254_driver_cuda.h):
# Fp elements as 4 x u64 LE in Montgomery form x 2^256 mod p; affine points as
# x[4] || y[4] || inf (9 x u64).
BN_P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47


def bn_fp_ref(x: int) -> bytes:
    return (x * (1 << 256) % BN_P).to_bytes(32, "little")


def bn_fp_from_ref(b: bytes) -> int:
    return int.from_bytes(b, "little") * pow(1 << 256, -1, BN_P) % BN_P


def bn_point_ref(P) -> bytes:
    """(x, y) plain ints or None (identity) -> 72-byte record."""
    if P is None:
        return bytes(64) + (1).to_bytes(8, "little")
    return bn_fp_ref(P[0]) + bn_fp_ref(P[1]) + bytes(8)


def bn_point_from_ref(b: bytes):
    if int.from_bytes(b[64:72], "little"):
        return None
    return bn_fp_from_ref(b[:32]), bn_fp_from_ref(b[32:64])


def fr_of_curve(curve: int) -> int:
    return 0 if curve == CURVE_BLS12_381 else 1


def fr_dot(fr: int, a: bytes, b: bytes) -> int:
    """sum a_i b_i mod r over two arrays of 32-byte LE values (O(n), C)."""
    assert len(a) == len(b) and len(a) % 32 == 0
    out = _buf(32)
    orc().orc_fr_dot(fr, a, b, len(a) // 32, out)
    return int.from_bytes(out.raw, "little")


def fr_eval(fr: int, coeffs: bytes, x: int) -> int | None:
    """sum c_i x^i (Horner, C)."""
    out = _buf(32)
    if orc().orc_fr_eval(fr, coeffs, len(coeffs) // 32, x.to_bytes(32, "little"), out):
        return None
    return int.from_bytes(out.raw, "little")


def fr_divide_linear(fr: int, coeffs: bytes, z: int) -> tuple[bytes, int] | None:
    """Synthetic division f = (X - z) q + y (C): (q's n - 1 coefficients, y)."""
    n = len(coeffs) // 32
    q, y = _buf(32 * max(n - 1, 1)), _buf(32)
    if orc().orc_fr_divide_linear(fr, coeffs, n, z.to_bytes(32, "little"), q, y):
        return None
    return q.raw[:32 * (n - 1)], int.from_bytes(y.raw, "little")


def fr_lagrange(fr: int, evals: bytes, tau: int) -> int | None:
    """f(tau) from the evaluations f(w^j) (barycentric formula, C)."""
    out = _buf(32)
    if orc().orc_fr_lagrange(fr, evals, len(evals) // 32, tau.to_bytes(32, "little"), out):
        return None
    return int.from_bytes(out.raw, "little")


def ref_available() -> bool:
    return os.path.exists(LIBREF)


def ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(LIBREF)
        _ref.gpukit_multi_pippenger_cpu.argtypes = [ctypes.c_uint32, _u8p, _u8p, _sz, _u8p]
    return _ref


def _buf(n):
    return ctypes.create_string_buffer(n)


# ---------------------------------------------------------------- oracle API
def field_op(field: int, op: int, a: bytes, b: bytes | None = None) -> bytes | None:
    nb = orc().orc_field_nbytes(field)
    out = _buf(nb)
    rc = orc().orc_field_op(field, op, a, b, out)
    return None if rc else out.raw


def field_modulus(field: int) -> int:
    nb = orc().orc_field_nbytes(field)
    out = _buf(nb)
    orc().orc_field_modulus(field, out)
    return int.from_bytes(out.raw, "little")


def mont_consts(field: int):
    nb = orc().orc_field_nbytes(field)
    r1, r2, pinv = _buf(nb), _buf(nb), ctypes.c_uint64()
    orc().orc_field_mont_consts(field, r1, r2, ctypes.byref(pinv))
    return int.from_bytes(r1.raw, "little"), int.from_bytes(r2.raw, "little"), pinv.value


def gen_scalars(curve: int, seed: int, n: int) -> bytes:
    out = _buf(32 * n)
    orc().orc_gen_scalars(curve, seed, n, out)
    return out.raw


def generator(curve: int) -> bytes:
    out = _buf(POINT_BYTES[curve])
    orc().orc_g1_generator(curve, out)
    return out.raw


def g1_mul_gen(curve: int, k: bytes) -> bytes:
    out = _buf(POINT_BYTES[curve])
    orc().orc_g1_mul_gen(curve, k, out)
    return out.raw


def g1_mul(curve: int, p: bytes, k: bytes) -> bytes:
    out = _buf(POINT_BYTES[curve])
    orc().orc_g1_mul(curve, p, k, out)
    return out.raw


def g1_add(curve: int, a: bytes, b: bytes) -> bytes:
    out = _buf(POINT_BYTES[curve])
    orc().orc_g1_add(curve, a, b, out)
    return out.raw


def g1_dbl(curve: int, a: bytes) -> bytes:
    out = _buf(POINT_BYTES[curve])
    orc().orc_g1_dbl(curve, a, out)
    return out.raw


def on_curve(curve: int, p: bytes) -> bool:
    return bool(orc().orc_g1_on_curve(curve, p))


def compress(curve: int, p: bytes) -> bytes:
    out = _buf(COMP_BYTES[curve])
    orc().orc_g1_compress(curve, p, out)
    return out.raw


def decompress(curve: int, c: bytes) -> bytes | None:
    out = _buf(POINT_BYTES[curve])
    rc = orc().orc_g1_decompress(curve, c, out)
    return None if rc else out.raw


def gen_points(curve: int, ks: bytes) -> bytes:
    n = len(ks) // 32
    out = _buf(POINT_BYTES[curve] * n)
    orc().orc_gen_points(curve, ks, n, out)
    return out.raw


def gen_points_progression(curve: int, k0: int, d: int, n: int) -> bytes:
    """(k0 + i*d) * G for i < n."""
    out = _buf(POINT_BYTES[curve] * n)
    orc().orc_gen_points_progression(curve, k0.to_bytes(32, "little"), d.to_bytes(32, "little"), n, out)
    return out.raw


def msm(curve: int, scalars: bytes, points: bytes, threads: int = 1) -> bytes:
    n = len(scalars) // 32
    out = _buf(POINT_BYTES[curve])
    if threads > 1:
        orc().orc_msm_mt(curve, scalars, points, n, out, threads)
    else:
        orc().orc_msm(curve, scalars, points, n, out)
    return out.raw


def msm_naive(curve: int, scalars: bytes, points: bytes) -> bytes:
    n = len(scalars) // 32
    out = _buf(POINT_BYTES[curve])
    orc().orc_msm_naive(curve, scalars, points, n, out)
    return out.raw


def ntt(fr: int, elems: bytes, inverse: bool = False) -> bytes | None:
    n = len(elems) // 32
    buf = ctypes.create_string_buffer(elems, len(elems))
    rc = orc().orc_ntt(fr, buf, n, 1 if inverse else 0)
    return None if rc else buf.raw[: len(elems)]


def dft_naive(fr: int, elems: bytes, inverse: bool = False) -> bytes | None:
    n = len(elems) // 32
    out = _buf(len(elems))
    rc = orc().orc_dft_naive(fr, elems, out, n, 1 if inverse else 0)
    return None if rc else out.raw


def root_of_unity(fr: int, n: int) -> int:
    out = _buf(32)
    if orc().orc_root_of_unity(fr, n, out):
        raise ValueError(n)
    return int.from_bytes(out.raw, "little")


def powers(curve: int, tau: bytes, n: int) -> bytes:
    out = _buf(32 * n)
    orc().orc_powers(curve, tau, n, out)
    return out.raw


def kzg_commit(curve: int, srs: bytes, evals: bytes) -> bytes | None:
    n = len(evals) // 32
    out = _buf(COMP_BYTES[curve])
    rc = orc().orc_kzg_commit(curve, srs, evals, n, out)
    return None if rc else out.raw


def kzg_commit_closed_form(curve: int, tau: bytes, evals: bytes) -> bytes | None:
    n = len(evals) // 32
    out = _buf(COMP_BYTES[curve])
    rc = orc().orc_kzg_commit_closed_form(curve, tau, evals, n, out)
    return None if rc else out.raw


# ------------------------------------------------------------- reference API
def ref_msm(curve: int, scalars: bytes, points: bytes) -> bytes:
    n = len(scalars) // 32
    out = _buf(POINT_BYTES[curve])
    rc = ref().gpukit_multi_pippenger_cpu(curve, scalars, points, n, out)
    assert rc == 0, rc
    return out.raw


BATCH_INV_BYTES = {0: 32, 1: 32, 2: 48}  # secp256k1, BN254, BLS12-381 base fields
_BATCH_INV_CPU = {0: "gpukit_batch_inv_secp256k1_fp_cpu", 1: "gpukit_batch_inv_bn254_fp_cpu",
                  2: "gpukit_batch_inv_bls12_381_fp_cpu"}


def ref_batch_inv(curve: int, data: bytes) -> tuple[int, bytes]:
    """The reference's CPU batch inversion (batch_inversion.cpp): (status, out)."""
    n = len(data) // BATCH_INV_BYTES[curve]
    out = ctypes.create_string_buffer(b"\xa5" * max(1, len(data)), max(1, len(data)))
    fn = getattr(ref(), _BATCH_INV_CPU[curve])
    fn.argtypes = [_u8p, _u8p, _sz]
    rc = fn(data, out, n)
    return rc, out.raw[: len(data)]


def ref_field_mul(field: int, a: bytes, b: bytes) -> bytes:
    nb = 48 if field == FIELD_BLS_FP else 32
    out = _buf(nb)
    ref().ref_field_mul(field, a, b, out)
    return out.raw


def ref_g1_add(curve: int, a: bytes, b: bytes) -> bytes:
    out = _buf(POINT_BYTES[curve])
    ref().ref_g1_add(curve, a, b, out)
    return out.raw


def ref_g1_dbl(curve: int, a: bytes) -> bytes:
    out = _buf(POINT_BYTES[curve])
    ref().ref_g1_dbl(curve, a, out)
    return out.raw


def ref_g1_mul(curve: int, p: bytes, k: bytes) -> bytes:
    out = _buf(POINT_BYTES[curve])
    ref().ref_g1_mul(curve, p, k, out)
    return out.raw


def ref_generator(curve: int) -> bytes:
    out = _buf(POINT_BYTES[curve])
    ref().ref_g1_generator(curve, out)
    return out.raw


# ---------------------------------------------------------------- Banderwagon
def bw_generator() -> bytes:
    out = _buf(64)
    orc().orc_bw_generator(out)
    return out.raw


def bw_check(p: bytes) -> bool:
    return bool(orc().orc_bw_check(p))


def bw_mul(p: bytes, k: int | bytes) -> bytes:
    kb = k if isinstance(k, bytes) else k.to_bytes(32, "little")
    out = _buf(64)
    orc().orc_bw_mul(p, kb, out)
    return out.raw


def bw_mul_gen(k: int | bytes) -> bytes:
    return bw_mul(bw_generator(), k)


def bw_add(a: bytes, b: bytes) -> bytes:
    out = _buf(64)
    orc().orc_bw_add(a, b, out)
    return out.raw


def bw_dbl(a: bytes) -> bytes:
    out = _buf(64)
    orc().orc_bw_dbl(a, out)
    return out.raw


def bw_compress(p: bytes) -> bytes:
    out = _buf(32)
    orc().orc_bw_compress(p, out)
    return out.raw


def bw_msm(scalars: bytes, points: bytes) -> tuple[int, bytes]:
    out = _buf(64)
    rc = orc().orc_bw_msm(scalars, points, len(scalars) // 32, out)
    return rc, out.raw


def bw_msm_naive(scalars: bytes, points: bytes) -> tuple[int, bytes]:
    out = _buf(64)
    rc = orc().orc_bw_msm_naive(scalars, points, len(scalars) // 32, out)
    return rc, out.raw


def bw_gen_points(ks: bytes) -> bytes:
    out = _buf(max(1, 2 * len(ks)))
    orc().orc_bw_gen_points(ks, len(ks) // 32, out)
    return out.raw[: 2 * len(ks)]


def ref_bw_msm(scalars: bytes, points: bytes) -> tuple[int, bytes]:
    """The reference's gpukit_multi_pippenger_cpu on curve 3: (status, 64 bytes)."""
    out = _buf(64)
    rc = ref().gpukit_multi_pippenger_cpu(CURVE_BANDERWAGON, scalars, points, len(scalars) // 32, out)
    return rc, out.raw
