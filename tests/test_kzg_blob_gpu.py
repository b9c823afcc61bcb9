"""EIP-4844 blob entries kzg_blob_to_commit / kzg_commit_to_proof
(include/kinet/crypto/kzg.h; the reference's c-abi/crypto.h:238-239) over a
synthetic trusted setup in c-kzg-4844's text format written from a known tau
(tests/kzg_setup.py; the mainnet file is not available here), through the C
ABI.  Checks:
  - the reference's fixtures (kzg/test/kzg_test.cpp:38-42, 71-81): the zero
    blob commits to infinity (0xC0 || 0), the all-ones blob to the generator
    (0x80 | x_G); both open to an infinity proof with y = 0 / 1;
  - random blobs against [f(tau)] G1, f interpolated on the bit-reversed
    domain (barycentric closed form in the C oracle);
  - proofs at z outside the domain and at domain points (the EIP-4844 special
    case) against y = f(z) and [(f(tau) - y) / (tau - z)] G1;
  - rejections: elements >= r, z >= r, setups with a point outside G1 or a bad
    encoding; loading from KINET_KZG_TRUSTED_SETUP on first use."""
import ctypes
import json
import os
import random

import pytest

import kzg_setup as K
import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu

TAU = 0x2A9B3C4D5E6F708192A3B4C5D6E7F8091A2B3C4D5E6F708192A3B4C5D6E7F80
R = M.BLS_R
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    import crypto_amd

    lib = crypto_amd.lib.require_device()
    lib.kzg_load_trusted_setup_file.argtypes = [ctypes.c_char_p]
    lib.kzg_free_trusted_setup.restype = None
    for name in ("kzg_blob_to_commit",):
        getattr(lib, name).argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.kzg_commit_to_proof.argtypes = [ctypes.c_char_p] * 4
    path = tmp_path_factory.mktemp("setup") / "trusted_setup.txt"
    print("writing a synthetic trusted setup (4096 + 4096 G1 points)", flush=True)
    path.write_text(K.setup_text(TAU, K.oracle_mul_gen_batch))
    assert lib.kzg_load_trusted_setup_file(str(path).encode()) == 0, crypto_amd.lib.last_error()
    yield lib, str(path)
    lib.kzg_free_trusted_setup()


def _commit(lib, blob):
    out = ctypes.create_string_buffer(48)
    rc = lib.kzg_blob_to_commit(blob, out)
    return rc, out.raw


def _proof(lib, blob, z_be):
    proof, y = ctypes.create_string_buffer(48), ctypes.create_string_buffer(32)
    rc = lib.kzg_commit_to_proof(blob, z_be, proof, y)
    return rc, proof.raw, int.from_bytes(y.raw, "big")


def _blob(seed):
    rnd = random.Random(seed)
    return b"".join(rnd.randrange(R).to_bytes(32, "big") for _ in range(K.N))


def _closed_point(v):
    return O.compress(2, O.g1_mul_gen(2, (v % R).to_bytes(32, "little")))


def test_reference_fixtures(setup):
    lib, _ = setup
    fx = json.load(open(os.path.join(GOLDEN, "bls_compressed.json")))
    print("blob fixtures: zero and all-ones blobs", flush=True)
    zero = bytes(131072)
    rc, c = _commit(lib, zero)
    assert rc == 0 and c == bytes([0xC0]) + bytes(47)
    z = bytearray(32)
    z[13] = 17  # kzg_test.cpp's arbitrary z
    rc, proof, y = _proof(lib, zero, bytes(z))
    assert rc == 0 and proof == bytes([0xC0]) + bytes(47) and y == 0
    ones = (1).to_bytes(32, "big") * K.N
    rc, c = _commit(lib, ones)
    gen = O.compress(2, O.generator(2))
    assert rc == 0 and c == gen
    assert gen[0] & 0x80 and gen[1:] == bytes.fromhex(
        "f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb")
    rc, proof, y = _proof(lib, ones, bytes(z))
    assert rc == 0 and proof == bytes([0xC0]) + bytes(47) and y == 1
    assert fx  # the compressed-point fixtures are also pinned on the CPU (test_oracle)


@pytest.mark.parametrize("seed", [1, 2])
def test_commit_and_proofs_closed_form(setup, seed):
    lib, _ = setup
    blob = _blob(seed)
    e_nat = K.natural_evals(blob)
    f_tau = O.fr_lagrange(0, e_nat, TAU)
    print(f"blob seed={seed}: commitment and proofs", flush=True)
    rc, c = _commit(lib, blob)
    assert rc == 0 and c == _closed_point(f_tau)
    rnd = random.Random(seed + 10)
    w = K.W
    # outside the domain, then two domain points: w^brp(5) (blob index 5) and 1 (index 0)
    zs = [rnd.randrange(R), rnd.randrange(R), pow(w, K.brp(5), R), 1]
    for z in zs:
        rc, proof, y = _proof(lib, blob, z.to_bytes(32, "big"))
        assert rc == 0
        fz = O.fr_lagrange(0, e_nat, z)
        if fz is None:  # z in the domain: y is the blob element there
            k = [j for j in range(K.N) if pow(w, K.brp(j), R) == z][0]
            fz = int.from_bytes(blob[32 * k:32 * k + 32], "big")
        assert y == fz
        assert proof == _closed_point((f_tau - fz) * pow(TAU - z, -1, R)), z


def test_rejections(setup, tmp_path):
    lib, _ = setup
    blob = bytearray(_blob(3))
    blob[32 * 7:32 * 8] = R.to_bytes(32, "big")  # an element equal to r
    assert _commit(lib, bytes(blob))[0] == -1
    assert _proof(lib, _blob(3), R.to_bytes(32, "big"))[0] == -1  # z = r
    assert _proof(lib, _blob(3), bytes([0xFF]) * 32)[0] == -1
    # a setup with a curve point outside G1 is refused, and the loaded one kept
    p, b = M.BLS_P, 4
    for x in range(1, 1000):
        yy = M.sqrt(2, (x ** 3 + b) % p)
        if yy is not None and M.mul(2, (x, yy), R) is not None:
            bad = M.compress(2, (x, yy))
            break
    path = tmp_path / "bad_g1.txt"
    print("blob setup rejections", flush=True)
    path.write_text(K.setup_text(TAU, K.oracle_mul_gen_batch, monomial=False, override={9: bad}))
    assert lib.kzg_load_trusted_setup_file(str(path).encode()) == -1
    path.write_text(K.setup_text(TAU, K.oracle_mul_gen_batch, monomial=False,
                                 override={9: bytes([0x80]) + bytes(46) + b"\x05"}))
    assert lib.kzg_load_trusted_setup_file(str(path).encode()) == -1
    rc, c = _commit(lib, bytes(131072))
    assert rc == 0 and c[0] == 0xC0


def test_setup_from_environment_on_first_use(setup, monkeypatch, tmp_path):
    lib, path = setup
    blob = _blob(4)
    rc, want = _commit(lib, blob)
    assert rc == 0
    lib.kzg_free_trusted_setup()
    monkeypatch.delenv("KINET_KZG_TRUSTED_SETUP", raising=False)
    assert _commit(lib, blob)[0] == -4  # no setup: CRYPTO_ERR_BACKEND
    monkeypatch.setenv("KINET_KZG_TRUSTED_SETUP", path)
    print("blob setup from the environment", flush=True)
    assert _commit(lib, blob) == (0, want)
    # a c-kzg 1.x file (no monomial points) loads as well
    lib.kzg_free_trusted_setup()
    monkeypatch.delenv("KINET_KZG_TRUSTED_SETUP")
    v1 = tmp_path / "trusted_setup_v1.txt"
    v1.write_text(K.setup_text(TAU, K.oracle_mul_gen_batch, monomial=False))
    assert lib.kzg_load_trusted_setup_file(str(v1).encode()) == 0
    assert _commit(lib, blob) == (0, want)
