"""Writes predictor_tables.bin, the generated crossing-predictor tables the
library embeds (predictor_tables.S), from the reference's generated Rust
sources:

  jpxl-encode-policy/src/quality_predictor_v2.rs     (knot model, qpv2-st-1)
  jpxl-encode-policy/src/quality_predictor_cases.rs  (case table, qpv2-cases-2)

Both files are trained data (generated by the reference's tooling), not
algorithm; the numbers are taken as printed (shortest round-trip decimal, so
float() recovers every double exactly).  Run in a checkout that holds the
reference:

  python jpxl_amd/csrc/extract_predictor_tables.py [/root/reference]

Layout (little-endian): the 8-byte magic "JPXLQPT1", seven u32 counts
(knot-model dimension D2, knots K, case dimension D, cases N, knot targets T,
neighbours, 0), then f64 arrays in this order:
  D2 centers, D2 scales, D2 x (lo, hi) z-range,
  K x (target, lower[D2+1], median[D2+1], candidate[D2+1], upper[D2+1],
       beta[D2+1], saturation[D2+1]),
  neighbour ratio, spread fallback, distance eps,
  D weights, D centers, D scales, T knot targets, D x (lo, hi) z-range,
  N x (z[D], ln_crossing[T], ln_beta[T]).
"""
from __future__ import annotations

import os
import re
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "JPXL/crates/jpxl-encode-policy/src"


def _numbers(text: str) -> list[float]:
    return [float(t) for t in re.findall(r"-?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?", text)]


def _const(src: str, name: str) -> str:
    m = re.search(r"pub const " + name + r"\s*:[^=]*=\s*(.*?);\n", src, re.S)
    if not m:
        raise SystemExit(f"constant {name} not found")
    return m.group(1)


def _array(src: str, name: str, n: int) -> list[float]:
    v = _numbers(_const(src, name))
    if len(v) != n:
        raise SystemExit(f"{name}: {len(v)} values, expected {n}")
    return v


def _records(src: str, const_name: str, struct_name: str) -> list[str]:
    start = src.index(f"pub const {const_name}:")
    body = src[start:]
    parts = body.split(f"{struct_name} {{")[1:]
    return parts


def _field(rec: str, name: str) -> list[float]:
    m = re.search(r"\b" + name + r":\s*\[(.*?)\]", rec, re.S)
    if not m:
        raise SystemExit(f"field {name} not found")
    return _numbers(m.group(1))


def main(root: str) -> None:
    v2 = open(os.path.join(root, SRC, "quality_predictor_v2.rs")).read()
    cs = open(os.path.join(root, SRC, "quality_predictor_cases.rs")).read()
    d2 = int(_const(v2, "QPV2_FEATURE_DIM"))
    knots = []
    for rec in _records(v2, "QPV2_KNOTS", "Qpv2Knot"):
        target = float(re.search(r"target:\s*([-\d.eE+]+)", rec).group(1))
        row = [target]
        for f in ("lower", "median", "candidate", "upper", "beta", "saturation"):
            w = _field(rec, f)
            assert len(w) == d2 + 1, (f, len(w))
            row += w
        knots.append(row)
    d = int(_const(cs, "QPV2_CASES_FEATURE_DIM"))
    t_targets = _array(cs, "QPV2_CASES_KNOT_TARGETS", 7)
    cases = []
    for rec in _records(cs, "QPV2_CASES", "Qpv2Case"):
        z, lc, lb = _field(rec, "z"), _field(rec, "ln_crossing"), _field(rec, "ln_beta")
        assert len(z) == d and len(lc) == len(t_targets) and len(lb) == len(t_targets)
        cases.append(z + lc + lb)
    out = bytearray(b"JPXLQPT1")
    out += struct.pack("<7I", d2, len(knots), d, len(cases), len(t_targets),
                       int(_const(cs, "QPV2_CASES_NEIGHBOURS")), 0)
    vals = []
    vals += _array(v2, "QPV2_FEATURE_CENTERS", d2)
    vals += _array(v2, "QPV2_FEATURE_SCALES", d2)
    vals += _array(v2, "QPV2_FEATURE_Z_RANGE", 2 * d2)
    for row in knots:
        vals += row
    vals += [float(_const(cs, "QPV2_CASES_NEIGHBOUR_RATIO")), float(_const(cs, "QPV2_CASES_SPREAD_FALLBACK")),
             float(_const(cs, "QPV2_CASES_DISTANCE_EPS"))]
    vals += _array(cs, "QPV2_CASES_FEATURE_WEIGHTS", d)
    vals += _array(cs, "QPV2_CASES_FEATURE_CENTERS", d)
    vals += _array(cs, "QPV2_CASES_FEATURE_SCALES", d)
    vals += t_targets
    vals += _array(cs, "QPV2_CASES_FEATURE_Z_RANGE", 2 * d)
    for row in cases:
        vals += row
    out += struct.pack(f"<{len(vals)}d", *vals)
    path = os.path.join(HERE, "predictor_tables.bin")
    with open(path, "wb") as f:
        f.write(out)
    print(f"{path}: {len(knots)} knots of dimension {d2}, {len(cases)} cases of dimension {d}, {len(out)} bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
