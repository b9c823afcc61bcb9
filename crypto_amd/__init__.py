"""MI355X-native KZG-commitment hot path: G1 MSM (BLS12-381, BN254) and the
scalar-field NTT, behind the reference's C ABI.

This package is a thin Python mirror of that C ABI (ctypes over
crypto_amd/lib/libkinet_crypto_hip.so).  There is no CPU fallback: if the HIP
library is not built or no HIP device is usable, calls raise.
"""
from .lib import (  # noqa: F401
    CURVE_BANDERWAGON,
    CURVE_BN254_G1,
    CURVE_BLS12_381_G1,
    FIELD_BLS12_381_FR,
    FIELD_BN254_FR,
    HipUnavailable,
    KinetError,
    KzgSrs,
    batch_inverse,
    bls12_381_g1_msm,
    kzg_commit_batch,
    library,
    library_path,
    multi_pippenger,
    ntt_fr,
    point_bytes,
    compressed_bytes,
)
