"""numpy wrappers around oracle/liboracle.so (the CPU restatement).

Test infrastructure only: the oracle is the checker, never the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_o = None


class QuantParams(ctypes.Structure):
    _fields_ = [
        ("global_scale", ctypes.c_uint32),
        ("hf_mul", ctypes.c_uint32),
        ("quant_lf", ctypes.c_uint32),
        ("x_qm", ctypes.c_int),
        ("b_qm", ctypes.c_int),
        ("truncate", ctypes.c_int),
        ("lambda_scale", ctypes.c_float),
        ("zero_token_bits", ctypes.c_float),
        ("aq_adj", ctypes.c_void_p),
        ("aq_refine_only", ctypes.c_int),
        ("aq_fine", ctypes.c_int),
        ("cover_legacy", ctypes.c_int),
    ]


class RenderParams(ctypes.Structure):
    _fields_ = [
        ("global_scale", ctypes.c_uint32),
        ("hf_mul", ctypes.c_uint32),
        ("quant_lf", ctypes.c_uint32),
        ("x_qm", ctypes.c_int),
        ("b_qm", ctypes.c_int),
        ("epf_iters", ctypes.c_int),
        ("gaborish", ctypes.c_int),
        ("sharpness", ctypes.c_int),
        ("bits", ctypes.c_int),
        ("aq_adj", ctypes.c_void_p),
        ("aq_refine_only", ctypes.c_int),
        ("aq_fine", ctypes.c_int),
    ]


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess

            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
        o = ctypes.CDLL(ORACLE_SO)
        f = ctypes.c_float
        for name in ["orc_cbrt_core", "orc_cbrt_metric", "orc_srgb_to_linear", "orc_linear_to_srgb"]:
            getattr(o, name).restype = f
            getattr(o, name).argtypes = [f]
        o.orc_scale_f.restype = f
        o.orc_scale_f.argtypes = [ctypes.c_int, ctypes.c_int]
        o.orc_remap.restype = ctypes.c_double
        o.orc_remap.argtypes = [ctypes.c_double]
        o.orc_ssimulacra2.restype = ctypes.c_double
        o.orc_quantize_frame.restype = ctypes.c_int
        o.orc_prepare_frame.restype = ctypes.c_int
        o.orc_depth_tables.restype = ctypes.c_int
        o.orc_aq_field.restype = ctypes.c_int
        o.orc_aq_mul_of.restype = ctypes.c_uint32
        o.orc_aq_mul_of.argtypes = [ctypes.c_uint32, ctypes.c_float]
        o.orc_aq_footprint_mul.restype = ctypes.c_uint32
        o.orc_aq_footprint_mul.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [ctypes.c_uint32, ctypes.c_int,
                                                                                   ctypes.c_int]
        o.orc_aq_setup.restype = ctypes.c_int
        o.orc_preanalysis.restype = ctypes.c_uint32
        _o = o
    return _o


def P(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def planes_ptr(arrs):
    t = ctypes.c_void_p * len(arrs)
    return t(*[a.ctypes.data for a in arrs])


def cbrt_core(v: float) -> float:
    return lib().orc_cbrt_core(v)


def cbrt_metric(v: float) -> float:
    return lib().orc_cbrt_metric(v)


def prepare_frame(rgb: np.ndarray, bits: int):
    """-> (xyb (3,H,W) f32, linear (3,H,W) f32, grayscale flag)."""
    h, w, _ = rgb.shape
    rgb = np.ascontiguousarray(rgb)
    xyb = np.zeros((3, h, w), np.float32)
    lin = np.zeros((3, h, w), np.float32)
    g = lib().orc_prepare_frame(w, h, bits, P(rgb), P(xyb[0]), P(xyb[1]), P(xyb[2]), P(lin[0]), P(lin[1]),
                                P(lin[2]))
    return xyb, lin, bool(g)


def dct2d(blocks: np.ndarray, inverse: bool = False) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, np.float32)
    out = np.empty_like(blocks)
    fn = lib().orc_dct2d_inv if inverse else lib().orc_dct2d_fwd
    n = blocks.shape[1]
    for i in range(blocks.shape[0]):
        fn(n, n, P(blocks[i]), P(out[i]))
    return out


def dct1d_naive(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    lib().orc_dct1d_naive(len(x), P(x), P(out), int(inverse))
    return out


def dct4_table(m: int) -> np.ndarray:
    out = np.empty((m, m), np.float32)
    lib().orc_dct4_table(m, P(out))
    return out


def dequant_matrix(side: int, channel: int) -> np.ndarray:
    out = np.empty((side, side), np.float32)
    lib().orc_dequant_matrix(side, channel, P(out))
    return out


def natural_order(side: int) -> np.ndarray:
    out = np.empty(side * side, np.uint32)
    lib().orc_natural_order(side, P(out))
    return out


def donor_weights(side: int) -> np.ndarray:
    out = np.empty(side * side, np.float32)
    lib().orc_donor_weights(side, P(out))
    return out


def depth_tables(bits: int):
    levels = 1 << bits
    thr = np.empty(levels - 1, np.float32)
    lut = np.empty(levels, np.float32)
    lib().orc_depth_tables(bits, P(thr), P(lut))
    return thr, lut


def forward_frame(xyb: np.ndarray, bmap: np.ndarray) -> np.ndarray:
    _, h, w = xyb.shape
    bh, bw = bmap.shape
    xyb = np.ascontiguousarray(xyb, np.float32)
    coeff = np.zeros((3, 8 * bh, 8 * bw), np.float32)
    bm = np.ascontiguousarray(bmap, np.uint8)
    lib().orc_forward_frame(w, h, planes_ptr([xyb[0], xyb[1], xyb[2]]), P(bm),
                            planes_ptr([coeff[0], coeff[1], coeff[2]]))
    return coeff


def quantize_frame(w, h, bmap, coeff, q: QuantParams, cfl_x=None, cfl_b=None, lf_factors=(0, 0)):
    bh, bw = bmap.shape
    th, tw = (bh + 7) // 8, (bw + 7) // 8
    cx = np.zeros((th, tw), np.int8) if cfl_x is None else np.ascontiguousarray(cfl_x, np.int8)
    cb = np.zeros((th, tw), np.int8) if cfl_b is None else np.ascontiguousarray(cfl_b, np.int8)
    quant = np.zeros((3, 8 * bh, 8 * bw), np.int32)
    lf = np.zeros((3, bh, bw), np.int32)
    coeff = np.ascontiguousarray(coeff, np.float32)
    bm = np.ascontiguousarray(bmap, np.uint8)
    rc = lib().orc_quantize_frame(w, h, P(bm), planes_ptr([coeff[0], coeff[1], coeff[2]]), ctypes.byref(q),
                                  P(cx), P(cb), int(lf_factors[0]), int(lf_factors[1]),
                                  planes_ptr([quant[0], quant[1], quant[2]]), planes_ptr([lf[0], lf[1], lf[2]]))
    if rc != 0:
        raise ValueError("oracle quantizer rejected a coefficient")
    return quant, lf


def render(w, h, bmap, quant, lf, r: RenderParams, cfl_x=None, cfl_b=None, lf_factors=(0, 0)):
    bh, bw = bmap.shape
    th, tw = (bh + 7) // 8, (bw + 7) // 8
    cx = np.zeros((th, tw), np.int8) if cfl_x is None else np.ascontiguousarray(cfl_x, np.int8)
    cb = np.zeros((th, tw), np.int8) if cfl_b is None else np.ascontiguousarray(cfl_b, np.int8)
    quant = np.ascontiguousarray(quant, np.int32)
    lf = np.ascontiguousarray(lf, np.int32)
    lin = np.zeros((3, h, w), np.float32)
    xyb = np.zeros((3, h, w), np.float32)
    bm = np.ascontiguousarray(bmap, np.uint8)
    lib().orc_render(w, h, P(bm), planes_ptr([quant[0], quant[1], quant[2]]), planes_ptr([lf[0], lf[1], lf[2]]),
                     P(cx), P(cb), int(lf_factors[0]), int(lf_factors[1]), ctypes.byref(r),
                     planes_ptr([lin[0], lin[1], lin[2]]), planes_ptr([xyb[0], xyb[1], xyb[2]]))
    return lin, xyb


def cover(xyb: np.ndarray, q: QuantParams) -> np.ndarray:
    """The hierarchical cover's block map (8/16/32 at varblock origins)."""
    _, h, w = xyb.shape
    xyb = np.ascontiguousarray(xyb, np.float32)
    bm = np.zeros(((h + 7) // 8, (w + 7) // 8), np.uint8)
    lib().orc_cover(w, h, planes_ptr([xyb[0], xyb[1], xyb[2]]), ctypes.byref(q), P(bm))
    return bm


def estimate_cfl(w: int, h: int, bmap: np.ndarray, coeff: np.ndarray, q: QuantParams):
    """-> (cfl_x tiles, cfl_b tiles, (x_factor_lf-128, b_factor_lf-128))."""
    bh, bw = bmap.shape
    th, tw = (bh + 7) // 8, (bw + 7) // 8
    cx = np.zeros((th, tw), np.int8)
    cb = np.zeros((th, tw), np.int8)
    lx, lb = ctypes.c_int(), ctypes.c_int()
    coeff = np.ascontiguousarray(coeff, np.float32)
    bm = np.ascontiguousarray(bmap, np.uint8)
    lib().orc_estimate_cfl(w, h, P(bm), planes_ptr([coeff[0], coeff[1], coeff[2]]), ctypes.byref(q), P(cx), P(cb),
                           ctypes.byref(lx), ctypes.byref(lb))
    return cx, cb, (lx.value, lb.value)


def epf_step(step: int, planes: np.ndarray, sigma: np.ndarray) -> np.ndarray:
    planes = np.ascontiguousarray(planes, np.float32)
    sigma = np.ascontiguousarray(sigma, np.float32)
    _, h, w = planes.shape
    out = np.zeros_like(planes)
    lib().orc_epf_step(step, w, h, planes_ptr([planes[0], planes[1], planes[2]]), P(sigma),
                       planes_ptr([out[0], out[1], out[2]]))
    return out


def gaborish_sharpen(planes: np.ndarray, iters: int = 12) -> np.ndarray:
    """The encoder's Jacobi Gaborish inverse of each plane."""
    planes = np.ascontiguousarray(planes, np.float32)
    out = np.empty_like(planes)
    h, w = planes.shape[1:]
    for c in range(planes.shape[0]):
        lib().orc_gaborish_sharpen(w, h, iters, P(planes[c]), P(out[c]))
    return out


def gaborish_forward_enc(planes: np.ndarray) -> np.ndarray:
    """The encoder's Gaborish forward convolution of each plane."""
    planes = np.ascontiguousarray(planes, np.float32)
    out = np.empty_like(planes)
    h, w = planes.shape[1:]
    for c in range(planes.shape[0]):
        lib().orc_gaborish_forward_enc(w, h, P(planes[c]), P(out[c]))
    return out


def gaborish(planes: np.ndarray) -> np.ndarray:
    planes = np.ascontiguousarray(planes, np.float32)
    _, h, w = planes.shape
    out = np.zeros_like(planes)
    for c in range(3):
        lib().orc_gaborish(w, h, P(planes[c]), P(out[c]))
    return out


def blur(a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    h, w = a.shape
    out = np.zeros_like(a)
    lib().orc_blur(w, h, P(a), None if b is None else P(np.ascontiguousarray(b, np.float32)), P(out))
    return out


def ssimulacra2(ref: np.ndarray, cand: np.ndarray):
    """-> (score, terms, raw_error); planar linear RGB (3, H, W)."""
    ref = np.ascontiguousarray(ref, np.float32)
    cand = np.ascontiguousarray(cand, np.float32)
    _, h, w = ref.shape
    terms = np.zeros(108, np.float64)
    ns = ctypes.c_int()
    raw = ctypes.c_double()
    s = lib().orc_ssimulacra2(w, h, planes_ptr([ref[0], ref[1], ref[2]]), planes_ptr([cand[0], cand[1], cand[2]]),
                              P(terms), ctypes.byref(ns), ctypes.byref(raw))
    return s, terms[: 18 * ns.value], raw.value


def remap(err: float) -> float:
    return lib().orc_remap(err)


def positive_xyb(rgb3) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_positive_xyb_pixel(ctypes.c_float(rgb3[0]), ctypes.c_float(rgb3[1]), ctypes.c_float(rgb3[2]), P(out))
    return out


def xyb_pixel(rgb3) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_xyb_pixel(ctypes.c_float(rgb3[0]), ctypes.c_float(rgb3[1]), ctypes.c_float(rgb3[2]), P(out))
    return out


def opsin_inverse(xyb3) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_opsin_inverse_pixel(ctypes.c_float(xyb3[0]), ctypes.c_float(xyb3[1]), ctypes.c_float(xyb3[2]), P(out))
    return out


def fixed_dct8_map(w: int, h: int) -> np.ndarray:
    return np.full(((h + 7) // 8, (w + 7) // 8), 8, np.uint8)


def mixed_map(w: int, h: int, seed: int) -> np.ndarray:
    """A random valid block map with DCT8/16/32 squares (tile-aligned)."""
    rng = np.random.default_rng(seed)
    bh, bw = (h + 7) // 8, (w + 7) // 8
    bm = np.zeros((bh, bw), np.uint8)
    taken = np.zeros((bh, bw), bool)
    for by in range(bh):
        for bx in range(bw):
            if taken[by, bx]:
                continue
            for side in (32, 16, 8):
                n = side // 8
                if side > 8 and rng.uniform() < 0.5:
                    continue
                if bx % n or by % n or bx + n > bw or by + n > bh:
                    continue
                if taken[by : by + n, bx : bx + n].any():
                    continue
                bm[by, bx] = side
                taken[by : by + n, bx : bx + n] = True
                break
    return bm


# --- adaptive quantization ------------------------------------------------
AQ_OFF, AQ_MASKING, AQ_UNIFORM, AQ_FINE_MASKING, AQ_FINE_UNIFORM, AQ_EDGE_REFINE = range(6)


def atlas_variance(xyb: np.ndarray) -> np.ndarray:
    """Per-atom sample variance, (3, bh, bw) f32."""
    _, h, w = xyb.shape
    xyb = np.ascontiguousarray(xyb, np.float32)
    out = np.zeros((3, (h + 7) // 8, (w + 7) // 8), np.float32)
    lib().orc_atlas_variance(w, h, planes_ptr([xyb[0], xyb[1], xyb[2]]), P(out))
    return out


def aq_field(xyb: np.ndarray, mode: int):
    """-> (adj (bh, bw) f32, active, fine): active is False for no field or
    a neutral one."""
    _, h, w = xyb.shape
    xyb = np.ascontiguousarray(xyb, np.float32)
    adj = np.zeros(((h + 7) // 8, (w + 7) // 8), np.float32)
    fine = ctypes.c_int()
    active = lib().orc_aq_field(w, h, planes_ptr([xyb[0], xyb[1], xyb[2]]), int(mode), P(adj), ctypes.byref(fine))
    return adj, bool(active), bool(fine.value)


def aq_setup(gs: int, quant_lf: int, hf_mul: int, fine: bool):
    """-> (global_scale, quant_lf, baseline hf_mul, refine_only) on the wire."""
    g, q, m = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    ro = lib().orc_aq_setup(gs, quant_lf, hf_mul, int(fine), ctypes.byref(g), ctypes.byref(q), ctypes.byref(m))
    return g.value, q.value, m.value, bool(ro)


def aq_mul_of(baseline: int, adj: float) -> int:
    return int(lib().orc_aq_mul_of(baseline, adj))


def aq_footprint_mul(adj: np.ndarray, bx: int, by: int, rows: int, cols: int, baseline: int, refine_only: bool,
                     fine: bool) -> int:
    adj = np.ascontiguousarray(adj, np.float32)
    bh, bw = adj.shape
    return int(lib().orc_aq_footprint_mul(P(adj), bw, bh, bx, by, rows, cols, baseline, int(refine_only),
                                          int(fine)))


def with_aq(params, adj: np.ndarray | None, refine_only: bool = False, fine: bool = False):
    """Attach an AQ field to QuantParams/RenderParams (kept alive on the object)."""
    if adj is None:
        params.aq_adj = None
        params._aq = None
    else:
        adj = np.ascontiguousarray(adj, np.float32)
        params._aq = adj
        params.aq_adj = adj.ctypes.data
    params.aq_refine_only = int(refine_only)
    params.aq_fine = int(fine)
    return params


def aq_block_muls(w: int, h: int, bmap: np.ndarray, q: QuantParams) -> np.ndarray:
    bm = np.ascontiguousarray(bmap, np.uint8)
    out = np.zeros(bm.shape, np.uint32)
    lib().orc_aq_block_muls(w, h, P(bm), ctypes.byref(q), P(out))
    return out


# ---- first-guess features of the quality search ------------------------------

PREANALYSIS_FIELDS = ["share", "log2var_mean", "log2var_std", "log2var_q10", "log2var_q90", "edge_mean",
                      "edge_q90", "flat_mean", "noise_mean", "noise_q90", "chroma_log2_mean", "chroma_log2_q90",
                      "orient_entropy_mean", "dir_dominance_mean", "axis_aligned_mean", "hetero_frac"]

TRANSFORM_FIELDS = ["blocks", "ln_ac_y_mean", "ln_ac_y_q50", "ln_ac_y_q90", "ln_ac_y_q99", "high_low_ratio",
                    "directional_asymmetry", "chroma_ac_ratio", "near_zero_frac_1e3", "near_zero_frac_1e2",
                    "dc_variance_y"]


def preanalysis(rgb: np.ndarray, bits: int):
    """preanalysis_srgb8/16 -> None (no whole 32x32 cell) or a dict of the
    PreanalysisFeatures fields (share as a (7,) f32 array), with 'cells'."""
    h, w, _ = rgb.shape
    rgb = np.ascontiguousarray(rgb, np.uint8 if bits == 8 else np.uint16)
    out = np.zeros(22, np.float32)
    n = lib().orc_preanalysis(w, h, bits, P(rgb), P(out))
    if n == 0:
        return None
    d = {"cells": int(n), "share": out[:7].copy()}
    for k, name in enumerate(PREANALYSIS_FIELDS[1:]):
        d[name] = np.float32(out[7 + k])
    return d


def transform_summary(xyb: np.ndarray):
    """quality_transform_summary -> (dict of the TransformFeatureSummary
    fields, the (512,) ln-energy histogram)."""
    _, h, w = xyb.shape
    xyb = np.ascontiguousarray(xyb, np.float32)
    out = np.zeros(11, np.float64)
    hist = np.zeros(512, np.uint32)
    lib().orc_transform_summary(w, h, planes_ptr([xyb[0], xyb[1], xyb[2]]), P(out), P(hist))
    d = {name: float(out[k]) for k, name in enumerate(TRANSFORM_FIELDS)}
    d["blocks"] = int(out[0])
    return d, hist


def floor_quantile(v: np.ndarray, q: float) -> np.float32:
    """quality_features.rs:141-152: the element at floor((n-1) q)."""
    if v.size == 0:
        return np.float32(0.0)
    s = np.sort(v.astype(np.float32))
    return s[int(np.floor((s.size - 1) * float(np.float32(q))))]


def source_features(xyb: np.ndarray, grayscale: bool):
    """source_features (quality_features.rs:173-213) from the atlas."""
    var = atlas_variance(xyb)
    luma = var[1].reshape(-1)
    chroma = (var[0] + var[2]).reshape(-1).astype(np.float32)
    q10, q50, q90 = (floor_quantile(luma, q) for q in (0.1, 0.5, 0.9))
    flat = np.float32(np.count_nonzero(luma < np.float32(2e-5))) / np.float32(luma.size)
    return {"width": xyb.shape[2], "height": xyb.shape[1], "grayscale": bool(grayscale),
            "luma_variance_q10": q10, "luma_variance_q50": q50, "luma_variance_q90": q90,
            "chroma_variance_q50": floor_quantile(chroma, 0.5), "flat_fraction": np.float32(flat),
            "edge_proxy": np.float32(q90 - q50)}
