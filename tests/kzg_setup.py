"""Synthetic EIP-4844 trusted setups in c-kzg-4844's text format, for tests.

The mainnet setup is not available here, so tests write one from a known tau:
    4096
    <n2>
    4096 lines: compressed L_i(tau) G1, i in natural order, where
                L_i(X) = w^i (X^4096 - 1) / (4096 (X - w^i)), w = 7^((r-1)/4096)
    n2 lines:   96-byte compressed G2 encodings (not used by the library)
    4096 lines: compressed tau^i G1 (the monomial points of c-kzg 2.x files)
Points k G come from a caller-supplied batch multiplier (the device's
fixed-base kernel on the GPU box, the oracle elsewhere); compression is the
oracle's.
"""
from __future__ import annotations

import oracle_lib as O
import pymath as M

N = 4096
R = M.BLS_R
W = pow(7, (R - 1) // N, R)
# compressed BLS12-381 G2 generator (ZCash encoding); the library reads G2
# lines for their count and flags only
G2_GEN = ("93e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e"
          "024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8")


def brp(i: int, bits: int = 12) -> int:
    return int(format(i, f"0{bits}b")[::-1], 2)


def lagrange_coeffs(tau: int) -> list[int]:
    """L_i(tau) for i < 4096, natural order."""
    tn = (pow(tau, N, R) - 1) * pow(N, -1, R) % R
    out, wi = [], 1
    for _ in range(N):
        out.append(wi * tn * pow(tau - wi, -1, R) % R)
        wi = wi * W % R
    return out


def setup_text(tau: int, mul_gen_batch, n2: int = 65, monomial: bool = True, override: dict | None = None) -> str:
    """mul_gen_batch(list of ints) -> concatenated 96-byte wire points k G.
    override: {index: 48-byte compressed point} replacing Lagrange lines."""
    lag = lagrange_coeffs(tau)
    wire = mul_gen_batch(lag)
    lines = [str(N), str(n2)]
    for i in range(N):
        c = O.compress(2, wire[96 * i:96 * (i + 1)])
        if override and i in override:
            c = override[i]
        lines.append(c.hex())
    lines += [G2_GEN] * n2
    if monomial:
        pw = [pow(tau, i, R) for i in range(N)]
        mw = mul_gen_batch(pw)
        lines += [O.compress(2, mw[96 * i:96 * (i + 1)]).hex() for i in range(N)]
    return "\n".join(lines) + "\n"


def oracle_mul_gen_batch(ks: list[int]) -> bytes:
    return O.gen_points(2, b"".join(k.to_bytes(32, "little") for k in ks))


def natural_evals(blob_be: bytes) -> bytes:
    """Blob (BE, bit-reversed domain order) -> LE evaluations in natural
    domain order, e[brp(j)] = blob_j."""
    e = [b""] * N
    for j in range(N):
        e[brp(j)] = blob_be[32 * j:32 * j + 32][::-1]
    return b"".join(e)
