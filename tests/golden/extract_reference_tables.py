"""Writes reference_tables.json: the numeric tables and constants of the
reference's VarDCT pixel path that the oracle restates, read mechanically from
the reference's source text (nothing is imported or run).

Usage: python tests/golden/extract_reference_tables.py [/root/reference]

The tests compare the oracle's tables with these values, so a transcription
slip in the oracle (a swapped row, a dropped digit) fails a test.
"""
from __future__ import annotations

import json
import os
import re
import sys

NUM = r"-?[0-9][0-9_]*(?:\.[0-9_]*)?(?:e-?[0-9]+)?"


def _num(tok: str) -> float:
    return float(tok.replace("_", ""))


def _block(text: str, start_pat: str, open_ch: str = "[", close_ch: str = "];") -> str:
    m = re.search(start_pat, text)
    if not m:
        raise SystemExit(f"pattern not found: {start_pat}")
    i = text.index("=", m.end() - 1) + 1
    j = text.index(close_ch, i)
    return text[i:j + 1]


def _numbers(s: str) -> list[float]:
    # drop comments, keep numeric literals
    s = re.sub(r"//[^\n]*", "", s)
    return [_num(t) for t in re.findall(NUM, s)]


def _rows(block: str) -> list[list[float]]:
    block = re.sub(r"//[^\n]*", "", block)
    inner = block.strip()[1:]
    rows = re.findall(r"\[([^\[\]]*)\]", inner)
    return [[_num(t) for t in re.findall(NUM, r)] for r in rows]


def main(root: str) -> None:
    crates = os.path.join(root, "JPXL", "crates")
    rd = lambda rel: open(os.path.join(crates, rel), encoding="utf-8").read()  # noqa: E731
    out: dict = {"source": "JPXL/crates (text of the .rs files named per entry)"}

    dq = rd("jpxl-core/src/dequant.rs")
    out["dequant_dct_params"] = {
        "file": "jpxl-core/src/dequant.rs",
        "8": _rows(_block(dq, r"pub const DCT8X8_PARAMS: \[\[f64; 6\]; 3\]")),
        "16": _rows(_block(dq, r"pub const DCT16X16_PARAMS: \[\[f64; 7\]; 3\]")),
        "32": _rows(_block(dq, r"pub const DCT32X32_PARAMS: \[\[f64; 8\]; 3\]")),
    }
    # Table I.6's parameter sets of the other transform types (the decoder's
    # side: rectangles, DCT2x2 and the 64-point DCTs of indices 11 and 12)
    out["dequant_other_params"] = {
        "file": "jpxl-core/src/dequant.rs",
        "16x8": _rows(_block(dq, r"pub const DCT16X8_PARAMS: \[\[f64; 7\]; 3\]")),
        "32x8": _rows(_block(dq, r"pub const DCT32X8_PARAMS: \[\[f64; 8\]; 3\]")),
        "16x32": _rows(_block(dq, r"pub const DCT16X32_PARAMS: \[\[f64; 8\]; 3\]")),
        "dct2": _rows(_block(dq, r"pub const DCT2_PARAMS: \[\[f64; 6\]; 3\]")),
        "large_bases": _rows(_block(dq, r"pub const LARGE_DCT_BASES: \[\[f64; 3\]; 6\]")),
        "seq": [_numbers(_block(dq, r"pub const SEQ_%s: \[f64; 7\]" % k)) for k in "ABC"],
    }

    s2 = rd("jpxl-perceptual/src/ssimulacra2.rs")
    weights = _numbers(_block(s2, r"const WEIGHTS: \[f64; 108\]"))
    assert len(weights) == 108, len(weights)
    m = re.search(r"pub fn remap\(error: f64\) -> f64 \{(.*?)\n\}", s2, re.S)
    remap_consts = [_num(t) for t in re.findall(r"[0-9][0-9_]*\.[0-9_]+(?:e-?[0-9]+)?", m.group(1))]
    out["ssimulacra2"] = {"file": "jpxl-perceptual/src/ssimulacra2.rs", "weights": weights,
                          "remap_literals": remap_consts}
    pool = rd("jpxl-perceptual/src/pool.rs")
    out["ssim_c2"] = _num(re.search(r"pub const SSIM_C2: f32 = (" + NUM + ");", pool).group(1))

    rs = rd("jpxl-core/src/restoration.rs")
    grab = lambda pat: _num(re.search(pat + r"\s*(" + NUM + r")", rs).group(1))  # noqa: E731
    out["epf"] = {
        "file": "jpxl-core/src/restoration.rs",
        "channel_scale": _numbers(_block(rs, r"pub const DEFAULT_EPF_CHANNEL_SCALE: \[f32; 3\]")),
        "sigma_skip": grab(r"pub const EPF_SIGMA_SKIP_THRESHOLD: f32 ="),
        "step_multiplier_base": grab(r"pub const EPF_STEP_MULTIPLIER_BASE: f32 ="),
        "quant_mul": grab(r"quant_mul:"),
        "pass0_sigma_scale": grab(r"pass0_sigma_scale:"),
        "pass2_sigma_scale": grab(r"pass2_sigma_scale:"),
        "border_sad_mul_expr": re.search(r"border_sad_mul: ([0-9. /]+),", rs).group(1).strip(),
    }

    col = rd("jpxl-core/src/color.rs")
    qb = re.search(r"pub const DEFAULT_QUANT_BIAS: \[f32; 3\] =\s*\[(.*?)\];", col, re.S).group(1)
    out["quant_bias_exprs"] = [e.strip().replace("_", "") for e in qb.split(",") if e.strip()]
    out["quant_bias_numerator"] = _num(re.search(r"pub const DEFAULT_QUANT_BIAS_NUMERATOR: f32 = (" + NUM + ");",
                                                 col).group(1))

    qp = rd("jpxl-encode-policy/src/quality_predictor.rs")
    entries = re.findall(r"InitialRungEntry \{\s*target: (" + NUM + r"),\s*luma_bucket: (\d+),\s*flat_bucket: (\d+),"
                         r"\s*global_scale: (\d+),\s*support: (\d+),\s*\}", qp)
    out["initial_rung_table"] = {
        "file": "jpxl-encode-policy/src/quality_predictor.rs",
        "luma_edges": _numbers(_block(qp, r"pub const LUMA_BUCKET_EDGES: \[f32; 4\]")),
        "flat_edges": _numbers(_block(qp, r"pub const FLAT_BUCKET_EDGES: \[f32; 2\]")),
        "fallback_log_fit": _numbers(_block(qp, r"pub const FALLBACK_LOG_FIT: \[f64; 4\]")),
        "entries": [[_num(t), int(lb), int(fb), int(gs), int(sp)] for t, lb, fb, gs, sp in entries],
    }
    feats = rd("jpxl-encode-policy/src/quality_features.rs")
    out["flat_variance"] = _num(re.search(r"pub const FLAT_VARIANCE: f32 = (" + NUM + ");", feats).group(1))

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
