"""GPU parity of the Fr NTT (ntt_fr_forward / ntt_fr_inverse) against the
oracle, plus size-independent properties at the full sizes."""
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu

FR = {0: M.BLS_R, 1: M.BN_R}


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.lib.require_device()
    return crypto_amd


def _rand(field, n, seed):
    rnd = random.Random(seed)
    return b"".join(rnd.randrange(FR[field]).to_bytes(32, "little") for _ in range(n))


@pytest.mark.parametrize("field", [0, 1])
@pytest.mark.parametrize("logn", [0, 1, 2, 3, 5, 8, 10, 11, 12, 13, 16])
def test_ntt_matches_oracle(ca, field, logn):
    n = 1 << logn
    x = _rand(field, n, 100 * field + logn)
    print(f"ntt field={field} n={n}", flush=True)
    assert ca.ntt_fr(field, x) == O.ntt(field, x)
    assert ca.ntt_fr(field, x, inverse=True) == O.ntt(field, x, inverse=True)


@pytest.mark.parametrize("field", [0, 1])
def test_ntt_naive_dft_and_structure(ca, field):
    n = 64
    x = _rand(field, n, 7)
    assert ca.ntt_fr(field, x) == O.dft_naive(field, x)
    delta = (1).to_bytes(32, "little") + bytes(32 * (n - 1))
    assert ca.ntt_fr(field, delta) == (1).to_bytes(32, "little") * n
    ones = (1).to_bytes(32, "little") * n
    assert ca.ntt_fr(field, ones) == n.to_bytes(32, "little") + bytes(32 * (n - 1))


@pytest.mark.parametrize("field,logn", [(0, 20), (1, 20), (0, 22), (1, 22), (0, 21), (1, 24)])
def test_ntt_roundtrip_large(ca, field, logn):
    """Full sizes: inverse(forward(x)) = x, and forward is linear."""
    import numpy as np

    n = 1 << logn
    from crypto_amd import inputs

    x = inputs.scalars_bytes(2 if field == 0 else 1, 31 + logn, n)
    print(f"roundtrip field={field} n=2^{logn}", flush=True)
    y = ca.ntt_fr(field, x)
    assert y != x
    assert ca.ntt_fr(field, y, inverse=True) == x
    # X_0 = sum x_j  (checks the first output against an independent sum)
    r = FR[field]
    xs = np.frombuffer(x, dtype="<u8").reshape(-1, 4)
    s0 = sum(inputs.to_int(xs)) % r
    assert int.from_bytes(y[:32], "little") == s0


@pytest.mark.parametrize("field", [0, 1])
def test_ntt_rejects_bad_input(ca, field):
    bad = (FR[field]).to_bytes(32, "little") + bytes(32 * 7)
    with pytest.raises(ca.KinetError) as e:
        ca.ntt_fr(field, bad)
    assert e.value.code == -1
    with pytest.raises(ca.KinetError):
        ca.ntt_fr(field, bytes(32 * 3))
    with pytest.raises(ca.KinetError):
        ca.ntt_fr(2, bytes(32 * 4))
