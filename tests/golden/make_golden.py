#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference.

Run in a container that holds the reference tree (default /root/reference),
after `make -C oracle ref` has built oracle/_ref/libref.so from the reference's
own sources.  The fixtures are data only: inputs and the reference's outputs.

  fr_bls_kat.json      the 10 BLS12-381 Fr add/mul vectors of
                       banderwagon/test/fp_kat.h (gnark-crypto generated)
  bn254_eip196.json    G, -G, 2G of bn254/test/bn254_kat_test.cpp:89-108
  bls_compressed.json  compressed infinity and generator, kzg/test/kzg_test.cpp:38-42,71-81
  field_golden.json    field products from the reference's CIOS mont_mul
  g1_golden.json       k*G, P+Q, 2P from the reference's Jacobian formulas
  msm_golden.json      gpukit_multi_pippenger_cpu outputs on seeded inputs
  msm_golden_secp.json the same for secp256k1 (curve 0), points (k0 + i d) G
                       from pymath, scalars below 2^255
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import sys
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402
import pymath as M  # noqa: E402

REF = os.environ.get("KINET_REFERENCE", "/root/reference")
MSM_SIZES = [1, 2, 8, 64, 256, 1024, 4096]
EXPLICIT_MAX = 64


def dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
        f.write("\n")


def parse_fp_kat():
    text = open(os.path.join(REF, "banderwagon/test/fp_kat.h")).read()
    out = []
    pat = re.compile(r'\{"([^"]+)",\s*"(add|mul)",\s*\{([^}]*)\},\s*\{([^}]*)\},\s*\{([^}]*)\}\}')
    for m in pat.finditer(text):
        vals = [bytes(int(x, 16) for x in g.replace(",", " ").split()) for g in m.groups()[2:]]
        out.append({"name": m.group(1), "op": m.group(2), "a": vals[0].hex(), "b": vals[1].hex(),
                    "r": vals[2].hex()})
    assert len(out) == 10, len(out)
    return out


def parse_eip196():
    text = open(os.path.join(REF, "bn254/test/bn254_kat_test.cpp")).read()
    consts = dict(re.findall(r'constexpr const char\* (G1_\w+) =\s*"([0-9a-f]+)"', text))
    return {"G_x": consts["G1_X"], "G_y": consts["G1_Y"], "negG_y": consts["G1_NEG_Y"],
            "twoG_x": consts["G1_2G_X"], "twoG_y": consts["G1_2G_Y"]}


def parse_bls_compressed():
    text = open(os.path.join(REF, "kzg/test/kzg_test.cpp")).read()
    m = re.search(r"G1_GEN_X\[48\]\s*=\s*\{([^}]*)\}", text)
    x = bytes(int(v, 16) for v in m.group(1).replace(",", " ").split())
    gen = bytearray(x)
    gen[0] |= 0x80
    inf = bytearray(48)
    inf[0] = 0xC0
    return {"infinity": bytes(inf).hex(), "generator": bytes(gen).hex()}


def _points_chunk(args):
    curve, ks = args
    return O.gen_points(curve, ks)


def gen_points_parallel(curve, ks, workers=8):
    n = len(ks) // 32
    step = max(1, (n + workers - 1) // workers)
    chunks = [(curve, ks[32 * i: 32 * min(n, i + step)]) for i in range(0, n, step)]
    with ProcessPoolExecutor(workers) as ex:
        return b"".join(ex.map(_points_chunk, chunks))


def msm_inputs(curve, n, seed):
    scalars = O.gen_scalars(curve, seed, n)
    ks = O.gen_scalars(curve, seed ^ 0x5EED, n)
    points = gen_points_parallel(curve, ks) if n >= 512 else O.gen_points(curve, ks)
    return scalars, ks, points


def msm_cases():
    cases = []
    for curve in (O.CURVE_BN254, O.CURVE_BLS12_381):
        pb = O.POINT_BYTES[curve]
        for n in MSM_SIZES:
            seed = 1000 * curve + n
            scalars, ks, points = msm_inputs(curve, n, seed)
            out = O.ref_msm(curve, scalars, points)
            case = {"curve": curve, "n": n, "seed": seed, "kind": "seeded",
                    "sha256_scalars": hashlib.sha256(scalars).hexdigest(),
                    "sha256_points": hashlib.sha256(points).hexdigest(),
                    "expected": out.hex()}
            if n <= EXPLICIT_MAX:
                case["scalars"] = scalars.hex()
                case["points"] = points.hex()
            cases.append(case)
            print("msm", curve, n, out.hex()[:16], flush=True)
        # edge cases with explicit inputs
        r = O.field_modulus(O.FIELD_BLS_FR if curve == O.CURVE_BLS12_381 else O.FIELD_BN_FR)
        scalars, ks, points = msm_inputs(curve, 16, 77 + curve)
        pts = [points[pb * i: pb * (i + 1)] for i in range(16)]
        scs = [scalars[32 * i: 32 * (i + 1)] for i in range(16)]
        g = O.generator(curve)
        neg = lambda p: p[: pb // 2] + (  # noqa: E731
            (O.field_modulus(O.FIELD_BLS_FP if curve == 2 else O.FIELD_BN_FP)
             - int.from_bytes(p[pb // 2:], "big")).to_bytes(pb // 2, "big"))
        edge = [
            ("identity_points", [bytes(pb)] * 3 + pts[:5], scs[:8]),
            ("zero_scalars", pts[:8], [bytes(32)] * 4 + scs[:4]),
            ("repeated_point", [pts[0]] * 8, scs[:8]),
            ("point_and_negation", [pts[1], neg(pts[1]), pts[2], neg(pts[2])], [scs[3]] * 4),
            ("scalar_r_minus_1", [g, pts[3]], [(r - 1).to_bytes(32, "little"), scs[5]]),
            ("scalar_one", pts[:4], [(1).to_bytes(32, "little")] * 4),
            ("max_window_digits", pts[:6], [((1 << 254) - 1).to_bytes(32, "little"),
                                             (int("80" * 31 + "00", 16)).to_bytes(32, "little"),
                                             (int("7f" * 31 + "3f", 16)).to_bytes(32, "little"),
                                             (2 ** 200 + 2 ** 100 + 1).to_bytes(32, "little"),
                                             ((1 << 200) - 1).to_bytes(32, "little"),
                                             (1 << 128).to_bytes(32, "little")]),
        ]
        for name, p, s in edge:
            pb_all, sb_all = b"".join(p), b"".join(s)
            out = O.ref_msm(curve, sb_all, pb_all)
            cases.append({"curve": curve, "n": len(s), "kind": "explicit", "name": name,
                          "scalars": sb_all.hex(), "points": pb_all.hex(), "expected": out.hex()})
    return cases


def secp_points(n, k0, d):
    """(k0 + i d) G for i < n on secp256k1, as 64-byte BE wire points."""
    P, D = M.mul(0, M.CURVES[0]["g"], k0), M.mul(0, M.CURVES[0]["g"], d)
    out = []
    for _ in range(n):
        out.append(M.to_wire(0, P))
        P = M.add(0, P, D)
    return b"".join(out)


def secp_scalars(n, seed):
    """n seeded 32-byte LE scalars below 2^255 (the reference drops a final
    window carry, which scalars below 2^255 never produce)."""
    import random

    rnd = random.Random(seed)
    return b"".join(rnd.getrandbits(255).to_bytes(32, "little") for _ in range(n))


def secp_msm_cases():
    cases = []
    for n in MSM_SIZES:
        seed, k0, d = 5000 + n, 0x5EC9 + n, 0xD1FF0000 + 7 * n
        scalars, points = secp_scalars(n, seed), secp_points(n, k0, d)
        out = O.ref_msm(0, scalars, points)
        case = {"curve": 0, "n": n, "seed": seed, "k0": k0, "d": d, "kind": "seeded",
                "sha256_scalars": hashlib.sha256(scalars).hexdigest(),
                "sha256_points": hashlib.sha256(points).hexdigest(), "expected": out.hex()}
        if n <= EXPLICIT_MAX:
            case["scalars"], case["points"] = scalars.hex(), points.hex()
        cases.append(case)
        print("msm secp256k1", n, out.hex()[:16], flush=True)
    pts = secp_points(8, 0xABCDEF, 0x1234567)
    pt = [pts[64 * i: 64 * (i + 1)] for i in range(8)]
    sc = [secp_scalars(8, 99)[32 * i: 32 * (i + 1)] for i in range(8)]
    neg = lambda w: w[:32] + ((M.SECP_P - int.from_bytes(w[32:], "big")) % M.SECP_P).to_bytes(32, "big")  # noqa: E731
    edge = [
        ("identity_points", [bytes(64)] * 3 + pt[:5], sc),
        ("zero_scalars", pt, [bytes(32)] * 4 + sc[:4]),
        ("repeated_point", [pt[0]] * 8, sc),
        ("point_and_negation", [pt[1], neg(pt[1]), pt[2], neg(pt[2])], [sc[3]] * 4),
        ("max_digits_below_2_255", pt[:4], [((1 << 255) - 1).to_bytes(32, "little"),
                                            int("7f" + "80" * 31, 16).to_bytes(32, "little"),
                                            (1 << 254).to_bytes(32, "little"), (1).to_bytes(32, "little")]),
    ]
    for name, p, s in edge:
        pb_all, sb_all = b"".join(p), b"".join(s)
        cases.append({"curve": 0, "n": len(s), "kind": "explicit", "name": name, "scalars": sb_all.hex(),
                      "points": pb_all.hex(), "expected": O.ref_msm(0, sb_all, pb_all).hex()})
    return cases


def field_cases():
    import random
    rnd = random.Random(20261203)
    cases = []
    for f in range(4):
        p = O.field_modulus(f)
        nb = 48 if f == O.FIELD_BLS_FP else 32
        vals = [0, 1, 2, p - 1, p - 2, (p - 1) // 2] + [rnd.randrange(p) for _ in range(26)]
        for i in range(0, len(vals), 2):
            for a, b in ((vals[i], vals[i + 1]), (vals[i + 1], rnd.randrange(p))):
                ab, bb = a.to_bytes(nb, "little"), b.to_bytes(nb, "little")
                cases.append({"field": f, "a": ab.hex(), "b": bb.hex(),
                              "mul": O.ref_field_mul(f, ab, bb).hex()})
    return cases


def g1_cases():
    import random
    rnd = random.Random(4242)
    cases = []
    for curve in (O.CURVE_BN254, O.CURVE_BLS12_381):
        g = O.ref_generator(curve)
        for t in range(6):
            k = rnd.getrandbits(252).to_bytes(32, "little")
            k2 = rnd.getrandbits(200).to_bytes(32, "little")
            p = O.ref_g1_mul(curve, g, k)
            q = O.ref_g1_mul(curve, g, k2)
            cases.append({"curve": curve, "k": k.hex(), "kG": p.hex(), "k2": k2.hex(), "k2G": q.hex(),
                          "sum": O.ref_g1_add(curve, p, q).hex(), "dbl": O.ref_g1_dbl(curve, p).hex()})
    return cases


def main():
    if not O.ref_available():
        sys.exit("build the reference library first: make -C oracle ref")
    dump("fr_bls_kat.json", parse_fp_kat())
    dump("bn254_eip196.json", parse_eip196())
    dump("bls_compressed.json", parse_bls_compressed())
    dump("field_golden.json", field_cases())
    dump("g1_golden.json", g1_cases())
    dump("msm_golden.json", msm_cases())
    dump("msm_golden_secp.json", secp_msm_cases())


if __name__ == "__main__":
    main()
