"""GPU parity of the KZG commit path (iNTT -> MSM -> compress) against the
oracle and the closed form [f(tau)] G."""
import random

import pytest

import oracle_lib as O
import pymath as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import crypto_amd

    crypto_amd.lib.require_device()
    return crypto_amd


def _evals(curve, n, seed):
    return O.gen_scalars(curve, seed, n)


@pytest.mark.parametrize("curve", [1, 2])
def test_srs_generate_matches_oracle(ca, curve):
    tau = int.from_bytes(O.gen_scalars(curve, 5, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, 64)
    want = O.gen_points(curve, O.powers(curve, tau.to_bytes(32, "little"), 64))
    assert srs.export() == want
    assert srs.export(10, 5) == want[10 * len(want) // 64: 15 * len(want) // 64]


@pytest.mark.parametrize("curve", [1, 2])
@pytest.mark.parametrize("n", [1, 2, 64, 1024, 4096])
def test_commit_matches_oracle(ca, curve, n):
    tau = int.from_bytes(O.gen_scalars(curve, 11, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, 4096)
    ev = _evals(curve, n, 1000 + n)
    print(f"kzg curve={curve} n={n}", flush=True)
    got = srs.commit(ev)
    assert got == O.kzg_commit(curve, srs.export(0, n), ev)
    assert got == O.kzg_commit_closed_form(curve, tau.to_bytes(32, "little"), ev)


@pytest.mark.parametrize("curve", [1, 2])
def test_commit_closed_form_large(ca, curve):
    n = 1 << 16
    tau = int.from_bytes(O.gen_scalars(curve, 12, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, n)
    ev = _evals(curve, n, 77)
    assert srs.commit(ev) == O.kzg_commit_closed_form(curve, tau.to_bytes(32, "little"), ev)


@pytest.mark.parametrize("curve", [1, 2])
def test_commit_special_polynomials(ca, curve):
    n = 256
    tau = int.from_bytes(O.gen_scalars(curve, 13, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, n)
    g = O.generator(curve)
    # constant polynomial 1: commitment = G; zero polynomial: infinity
    assert srs.commit((1).to_bytes(32, "little") * n) == O.compress(curve, g)
    assert srs.commit(bytes(32 * n)) == O.compress(curve, bytes(len(g)))


@pytest.mark.parametrize("curve", [1, 2])
def test_batch_and_msm(ca, curve):
    tau = int.from_bytes(O.gen_scalars(curve, 14, 1), "little")
    a = ca.KzgSrs.generate(curve, tau, 1024)
    b = ca.KzgSrs.generate(curve, tau + 1, 512)
    evs = [_evals(curve, 1024, 1), _evals(curve, 512, 2), _evals(curve, 128, 3)]
    got = ca.kzg_commit_batch([a, b, a], evs)
    assert got == [a.commit(evs[0]), b.commit(evs[1]), a.commit(evs[2])]
    sc = O.gen_scalars(curve, 15, 300)
    assert a.msm(sc) == O.msm(curve, sc, a.export(0, 300))


def _non_g1_point(curve):
    """A point of E(Fp) outside G1 (BLS12-381 has a large cofactor)."""
    p, b = M.CURVES[curve]["p"], M.CURVES[curve]["b"]
    for x in range(1, 1000):
        y = M.sqrt(curve, (x * x * x + b) % p)
        if y is not None and M.mul(curve, (x, y), M.CURVES[curve]["r"]) is not None:
            return (x, y)
    raise AssertionError("no point found")


@pytest.mark.parametrize("curve", [1, 2])
def test_batch_lone_job(ca, curve):
    """A batch commit of a single polynomial of at least 4096 evaluations over
    a table-free SRS in G1 runs the single-MSM pipeline with the GLV split:
    the same commitment as the single commit and the closed form, at 4096 and
    2^13 evaluations; a lone smaller job keeps the batched form."""
    tau = int.from_bytes(O.gen_scalars(curve, 16, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, 1 << 13)
    for n in (1024, 4096, 1 << 13):
        ev = _evals(curve, n, 300 + n)
        print(f"kzg lone batch job curve={curve} n={n}", flush=True)
        want = O.kzg_commit_closed_form(curve, tau.to_bytes(32, "little"), ev)
        assert ca.kzg_commit_batch([srs], [ev]) == [want]
        assert srs.commit(ev) == want


def test_batch_lone_job_srs_outside_g1(ca):
    """kzg_srs_create accepts on-curve points outside G1 and then keeps MSMs
    against that SRS off the GLV split.  A lone batch job of 4096 evaluations
    (over a G1 SRS the GLV pipeline) equals the single commit and the oracle's
    iNTT + MSM over the same points, the non-G1 point among them."""
    curve, n = 2, 4096
    pb = O.POINT_BYTES[curve]
    pts = bytearray(O.gen_points_progression(curve, 0x5151, 0x77, n))
    pts[7 * pb:8 * pb] = M.to_wire(curve, _non_g1_point(curve))
    srs = ca.KzgSrs.from_points(curve, bytes(pts))
    ev = _evals(curve, n, 4242)
    want = O.kzg_commit(curve, bytes(pts), ev)
    print("kzg lone batch job over a non-G1 SRS", flush=True)
    assert srs.commit(ev) == want
    assert ca.kzg_commit_batch([srs], [ev]) == [want]


@pytest.mark.parametrize("curve", [1, 2])
def test_commit_rejects_bad_input(ca, curve):
    srs = ca.KzgSrs.generate(curve, 3, 64)
    r = M.CURVES[curve]["r"]
    bad = r.to_bytes(32, "little") + bytes(32 * 63)
    for ev in (bad, bytes(32 * 3), bytes(32 * 128)):
        with pytest.raises(ca.KinetError) as e:
            srs.commit(ev)
        assert e.value.code == -1
    with pytest.raises(ca.KinetError):
        ca.KzgSrs.generate(curve, r, 8)


@pytest.mark.parametrize("curve", [1, 2])
def test_commit_coeffs_matches_evals_path(ca, curve):
    import ctypes

    n = 512
    srs = ca.KzgSrs.generate(curve, 99, n)
    ev = _evals(curve, n, 5)
    coeffs = O.ntt(0 if curve == 2 else 1, ev, inverse=True)
    out = ctypes.create_string_buffer(ca.compressed_bytes(curve))
    lib = ca.library()
    lib.kzg_commit_coeffs.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p]
    assert lib.kzg_commit_coeffs(srs.handle, coeffs, n, out) == 0
    assert out.raw == srs.commit(ev)


@pytest.mark.parametrize("curve", [1, 2])
def test_precomputed_srs_commits_unchanged(ca, curve):
    """An SRS with a window table (kzg_srs_precompute) gives the same
    commitments, MSMs and coefficient commits, for full size and prefixes."""
    n = 1 << 13
    tau = int.from_bytes(O.gen_scalars(curve, 21, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, n)
    evs = {m: _evals(curve, m, 300 + m) for m in (1, 16, 4096, n)}
    before = {m: srs.commit(e) for m, e in evs.items()}
    sc = O.gen_scalars(curve, 7, 5000)
    msm_before = srs.msm(sc)
    coeffs = O.gen_scalars(curve, 8, n - 3)
    open_before = [srs.open_coeffs(coeffs, z) for z in (5, 1)]
    nbytes = srs.precompute()
    assert nbytes > 0
    for m, e in evs.items():
        assert srs.commit(e) == before[m], m
    assert srs.msm(sc) == msm_before
    assert [srs.open_coeffs(coeffs, z) for z in (5, 1)] == open_before
    assert before[n] == O.kzg_commit_closed_form(curve, tau.to_bytes(32, "little"), evs[n])
    assert srs.precompute(9) > 0  # rebuilt with other window bits
    assert srs.commit(evs[4096]) == before[4096]
    with pytest.raises(ca.KinetError):
        srs.precompute(4)


def _eval_at(curve, fs, x):
    """f(x) for f given by its values on the NTT domain (natural order), by the
    barycentric formula (x outside the domain)."""
    r = M.CURVES[curve]["r"]
    n = len(fs)
    w = pow(7 if curve == 2 else 5, (r - 1) // n, r)
    acc, wi = 0, 1
    for f in fs:
        acc = (acc + f * wi * pow((x - wi) % r, -1, r)) % r
        wi = wi * w % r
    return (pow(x, n, r) - 1) * pow(n, -1, r) % r * acc % r


@pytest.mark.parametrize("curve", [1, 2])
@pytest.mark.parametrize("n", [2, 16, 4096])
def test_open_matches_closed_form(ca, curve, n):
    """kzg_open: y = f(z) and the proof [q(tau)] G, q = (f - y) / (X - z), against
    the closed form computed in Python from the evaluations (barycentric
    formula), for random and special z."""
    r = M.CURVES[curve]["r"]
    tau = int.from_bytes(O.gen_scalars(curve, 21, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, max(n, 64))
    ev = _evals(curve, n, 40 + n)
    fs = [int.from_bytes(ev[32 * i:32 * i + 32], "little") for i in range(n)]
    rnd = random.Random(n)
    for z in (rnd.randrange(r), r - 2, 2):  # -1 = w^(n/2) is a domain point; -2 and 2 are not
        print(f"kzg open curve={curve} n={n} z={z % 1000}", flush=True)
        proof, y = srs.open(ev, z)
        assert y == _eval_at(curve, fs, z)
        q_tau = (_eval_at(curve, fs, tau) - y) * pow((tau - z) % r, -1, r) % r
        assert proof == O.compress(curve, O.g1_mul_gen(curve, q_tau.to_bytes(32, "little")))


@pytest.mark.parametrize("curve", [1, 2])
def test_open_rejects_domain_points_and_bad_inputs(ca, curve):
    import crypto_amd

    r = M.CURVES[curve]["r"]
    n = 16
    srs = ca.KzgSrs.generate(curve, 12345, 64)
    ev = _evals(curve, n, 77)
    w = pow(7 if curve == 2 else 5, (r - 1) // n, r)
    for z in (1, w, pow(w, 5, r), r, r + 1):  # domain points, then non-canonical z
        with pytest.raises(crypto_amd.KinetError):
            srs.open(ev, z % (1 << 256))
    with pytest.raises(crypto_amd.KinetError):
        srs.open(ev[:32 * 12], 9)  # not a power of two


def _identity_compressed(curve):
    return O.compress(curve, bytes(96 if curve == 2 else 64))


@pytest.mark.parametrize("curve", [1, 2])
@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 1000, 1025, 4096])
def test_open_coeffs_matches_oracle(ca, curve, n):
    """kzg_open_coeffs against the oracle's synthetic division and its MSM of
    the quotient over the same SRS points, for random z, 0, 1, -1 and a
    domain point; n crosses the 32-element segments of the device scan."""
    r = M.CURVES[curve]["r"]
    fr = O.fr_of_curve(curve)
    tau = int.from_bytes(O.gen_scalars(curve, 31, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, 4096)
    pts = srs.export(0, max(n - 1, 1))
    c = O.gen_scalars(curve, 500 + n, n)
    rnd = random.Random(n)
    for z in (rnd.randrange(r), 0, 1, r - 1, O.root_of_unity(fr, 64)):
        print(f"kzg open_coeffs curve={curve} n={n} z={z % 1000}", flush=True)
        proof, y = srs.open_coeffs(c, z)
        q, y_want = O.fr_divide_linear(fr, c, z)
        assert y == y_want
        want = O.compress(curve, O.msm(curve, q, pts)) if n > 1 else _identity_compressed(curve)
        assert proof == want


@pytest.mark.parametrize("curve", [1, 2])
def test_open_coeffs_closed_form_large(ca, curve):
    """At n = 2^20 + 3 coefficients (five scan levels, ragged at each): y = f(z)
    by Horner and the proof equals [(f(tau) - y) / (tau - z)] G."""
    r = M.CURVES[curve]["r"]
    fr = O.fr_of_curve(curve)
    n = (1 << 20) + 3
    tau = int.from_bytes(O.gen_scalars(curve, 32, 1), "little")
    srs = ca.KzgSrs.generate(curve, tau, n)
    c = O.gen_scalars(curve, 600, n)
    z = random.Random(7).randrange(r)
    print(f"kzg open_coeffs closed form curve={curve} n={n}", flush=True)
    proof, y = srs.open_coeffs(c, z)
    assert y == O.fr_eval(fr, c, z)
    q_tau = (O.fr_eval(fr, c, tau) - y) * pow((tau - z) % r, -1, r) % r
    assert proof == O.compress(curve, O.g1_mul_gen(curve, q_tau.to_bytes(32, "little")))
    assert srs.commit_coeffs(c) == O.compress(curve, O.g1_mul_gen(curve, O.fr_eval(fr, c, tau).to_bytes(32, "little")))


@pytest.mark.parametrize("curve", [1, 2])
def test_open_coeffs_matches_evaluation_form(ca, curve):
    """The two openings agree where both apply: kzg_open on f's evaluations
    (the forward NTT of its coefficients) and kzg_open_coeffs on the
    coefficients, z outside the domain."""
    fr = O.fr_of_curve(curve)
    n = 256
    srs = ca.KzgSrs.generate(curve, 4242, n)
    c = O.gen_scalars(curve, 700, n)
    ev = O.ntt(fr, c)
    for z in (2, 123456789):
        assert srs.open_coeffs(c, z) == srs.open(ev, z)


@pytest.mark.parametrize("curve", [1, 2])
def test_open_coeffs_refusals(ca, curve):
    r = M.CURVES[curve]["r"]
    srs = ca.KzgSrs.generate(curve, 77, 16)
    c = O.gen_scalars(curve, 800, 16)
    for bad_c, z in ((b"", 3), (c + c[:32], 3), (c, r), (c[:32 * 15] + r.to_bytes(32, "little"), 3)):
        with pytest.raises(ca.KinetError) as e:
            srs.open_coeffs(bad_c, z)
        assert e.value.code == -1
