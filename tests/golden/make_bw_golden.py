#!/usr/bin/env python3
"""Writes the Banderwagon fixtures under tests/golden/ from the reference.

Run in a container that holds the reference tree (default /root/reference),
after `make -C oracle ref` has built oracle/_ref/libref.so (its Banderwagon
branch compiles banderwagon/cpp/*.cpp).  The fixtures are data only.

  bw_kat.json          the known-answer tables the reference's own tests hold:
                         enc       16 scalars s and the compressed [s]G
                                   (banderwagon/test/element_kat.h, generated by
                                   gnark-crypto through the Go reference)
                         bad       16 compressed encodings that must be rejected
                                   (same file); each holds an x that fails the
                                   subgroup test
                         msm       1000 (seed, n, compressed MSM result) triples
                                   (banderwagon/test/multiexp_kat.h); the inputs
                                   are rebuilt from the seed by the xorshift64*
                                   stream of banderwagon/test/multiexp_kat_test.cpp:41-110
  msm_golden_bw.json   gpukit_multi_pippenger_cpu(curve 3) outputs (status and
                       64-byte affine x || y) on seeded inputs and edge cases:
                       identity and two-torsion points, points off the curve,
                       outside the subgroup or non-canonical, scalars at and
                       above the group order
"""
from __future__ import annotations

import hashlib
import json
import os
import random
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

REF = os.environ.get("KINET_REFERENCE", "/root/reference")
Q = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001  # base field
R_BW = 0x1CFB69D4CA675F520CCE760202687600FF8F87007419047174FD06B52876E7E1  # group order
A = Q - 5
D = 0x6389C12633C267CBC66E3BF86BE3B6D8CB66677177E54F92B369F2F5188D58E7
MSM_SIZES = [1, 2, 3, 8, 64, 256, 1024, 4096]
EXPLICIT_MAX = 64


def dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
        f.write("\n")


def _bytes(group: str) -> bytes:
    return bytes(int(x, 16) for x in group.replace(",", " ").split())


def parse_element_kat():
    text = open(os.path.join(REF, "banderwagon/test/element_kat.h")).read()
    enc = [{"name": m.group(1), "scalar_le": _bytes(m.group(2)).hex(), "enc_be": _bytes(m.group(3)).hex()}
           for m in re.finditer(r'\{"(enc_s\d+)",\s*\{([^}]*)\},\s*\{([^}]*)\}\}', text)]
    bad = [{"name": m.group(1), "enc_be": _bytes(m.group(2)).hex()}
           for m in re.finditer(r'\{"(bad_\w+)",\s*\{([^}]*)\}\}', text)]
    assert len(enc) == 16 and len(bad) == 16, (len(enc), len(bad))
    return enc, bad


def parse_msm_kat():
    text = open(os.path.join(REF, "banderwagon/test/multiexp_kat.h")).read()
    out = [{"name": m.group(1), "seed": int(m.group(2), 16), "n": int(m.group(3)), "enc_be": _bytes(m.group(4)).hex()}
           for m in re.finditer(r'\{"(msm_\w+)",\s*0x([0-9a-fA-F]+),\s*(\d+),\s*\{([^}]*)\}\}', text)]
    assert len(out) == 1000, len(out)
    return out


def sqrt_q(a: int) -> int | None:
    """Tonelli-Shanks square root in the base field (2-adicity 32)."""
    a %= Q
    if a == 0:
        return 0
    if pow(a, (Q - 1) // 2, Q) != 1:
        return None
    s, t = 32, (Q - 1) >> 32
    z = 7  # a non-residue (the field's multiplicative generator)
    m, c, x, b = s, pow(z, t, Q), pow(a, (t + 1) // 2, Q), pow(a, t, Q)
    while b != 1:
        i, b2 = 0, b
        while b2 != 1:
            b2, i = b2 * b2 % Q, i + 1
        f = pow(c, 1 << (m - i - 1), Q)
        m, c, x, b = i, f * f % Q, x * f % Q, b * f * f % Q
    return x


def y_of(x: int) -> int | None:
    """A y with (x, y) on the curve, or None."""
    den = (D * x * x - 1) % Q
    if den == 0:
        return None
    return sqrt_q((A * x * x - 1) * pow(den, Q - 2, Q))


def wire(x: int, y: int) -> bytes:
    return x.to_bytes(32, "big") + y.to_bytes(32, "big")


def seeded_inputs(n: int, seed: int) -> tuple[bytes, bytes]:
    rnd = random.Random(seed)
    scalars = b"".join(rnd.randrange(R_BW).to_bytes(32, "little") for _ in range(n))
    ks = b"".join(rnd.getrandbits(253).to_bytes(32, "little") for _ in range(n))
    return scalars, O.bw_gen_points(ks)


def msm_cases(bad_kats):
    cases = []
    for n in MSM_SIZES:
        seed = 3000 + n
        scalars, points = seeded_inputs(n, seed)
        rc, out = O.ref_bw_msm(scalars, points)
        case = {"n": n, "seed": seed, "kind": "seeded", "status": rc, "expected": out.hex(),
                "sha256_scalars": hashlib.sha256(scalars).hexdigest(),
                "sha256_points": hashlib.sha256(points).hexdigest()}
        if n <= EXPLICIT_MAX:
            case["scalars"], case["points"] = scalars.hex(), points.hex()
        cases.append(case)
        print("bw msm", n, rc, out.hex()[:16], flush=True)
    scalars, points = seeded_inputs(16, 77)
    pts = [points[64 * i: 64 * (i + 1)] for i in range(16)]
    scs = [scalars[32 * i: 32 * (i + 1)] for i in range(16)]
    neg = lambda p: ((Q - int.from_bytes(p[:32], "big")) % Q).to_bytes(32, "big") + p[32:]  # noqa: E731
    two_torsion = wire(0, Q - 1)  # (0, -1): on the curve, in the checked subgroup, the group's identity
    # a curve point outside the subgroup: an x of the reference's bad-subgroup
    # encodings whose y exists
    outside = None
    for b in bad_kats:
        x = int(b["enc_be"], 16)
        y = y_of(x)
        if y is not None:
            outside = wire(x, y)
            break
    assert outside is not None
    px = int.from_bytes(pts[0][:32], "big")
    py = int.from_bytes(pts[0][32:], "big")
    edge = [
        ("identity_points", [bytes(64)] * 3 + pts[:5], scs[:8]),
        ("identity_encoded", [wire(0, 1), pts[1]], scs[:2]),
        ("two_torsion_point", [two_torsion, pts[2], two_torsion], scs[:3]),
        ("negated_representative", [pts[3], wire((Q - px) % Q, (Q - py) % Q)], [scs[4], scs[4]]),
        ("zero_scalars", pts[:8], [bytes(32)] * 4 + scs[:4]),
        ("repeated_point", [pts[0]] * 8, scs[:8]),
        ("point_and_negation", [pts[1], neg(pts[1]), pts[2], neg(pts[2])], [scs[3]] * 4),
        ("scalar_order_minus_1", [pts[5], pts[6]], [(R_BW - 1).to_bytes(32, "little"), scs[5]]),
        ("scalar_one", pts[:4], [(1).to_bytes(32, "little")] * 4),
        ("max_window_digits", pts[:4], [(int("80" * 31, 16)).to_bytes(32, "little"),
                                         (int("7f" * 31, 16)).to_bytes(32, "little"),
                                         ((1 << 252) - 1).to_bytes(32, "little"),
                                         (1 << 200).to_bytes(32, "little")]),
        # rejected inputs: every one returns GPUKIT_ERR_BAD_SIZE
        ("reject_scalar_at_order", pts[:2], [scs[0], R_BW.to_bytes(32, "little")]),
        ("reject_scalar_above_order", pts[:2], [((1 << 256) - 1).to_bytes(32, "little"), scs[0]]),
        ("reject_off_curve", [pts[0], wire(px, (py + 1) % Q)], scs[:2]),
        ("reject_noncanonical_x", [wire(px + Q, py) if px + Q < (1 << 256) else wire(Q, py), pts[1]], scs[:2]),
        ("reject_noncanonical_y", [pts[0], wire(int.from_bytes(pts[1][:32], "big"), Q)], scs[:2]),
        ("reject_outside_subgroup", [pts[0], outside], scs[:2]),
        ("reject_y_zero", [wire(1, 0), pts[1]], scs[:2]),
    ]
    for name, p, s in edge:
        pb_all, sb_all = b"".join(p), b"".join(s)
        rc, out = O.ref_bw_msm(sb_all, pb_all)
        cases.append({"n": len(s), "kind": "explicit", "name": name, "status": rc, "scalars": sb_all.hex(),
                      "points": pb_all.hex(), "expected": out.hex() if rc == 0 else None})
        print("bw edge", name, rc, flush=True)
    return cases


def main():
    if not O.ref_available():
        sys.exit("build the reference library first: make -C oracle ref")
    enc, bad = parse_element_kat()
    dump("bw_kat.json", {"enc": enc, "bad": bad, "msm": parse_msm_kat()})
    dump("msm_golden_bw.json", msm_cases(bad))


if __name__ == "__main__":
    main()
