"""ctypes binding of libjpxlg.so (the C ABI declared in include/jpxlg.h).

The shared library is built in-tree (``jpxl_amd/libjpxlg.so``) by
``__graft_entry__.build()``.  Loading fails loudly when it is missing: there
is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjpxlg.so")

OK = 0
E_INVALID = -1
E_UNSUPPORTED = -2
E_OOM = -3
E_DEVICE = -4
E_TOO_SMALL = -5
E_RANGE = -6

BUF_XYB = 0
BUF_COEFF = 1
BUF_RENDER_XYB = 2
BUF_LINEAR = 3
BUF_SIGMA = 4
DCT_EXACT = 0
DCT_FP32 = 1
DCT_BF16X3 = 2
BUF_REF_XYB0 = 5
BUF_AQ_FIELD = 6

COVER_FIXED_DCT8 = 0
COVER_HIERARCHICAL = 1
COVER_HIERARCHICAL_LEGACY = 2
QUALITY_STATUS = ["met", "met_adjacent_rungs", "met_work_cap", "saturated_top", "saturated_floor",
                  "under_target_work_cap", "rescued_fresh_structure"]
AQ_OFF, AQ_MASKING, AQ_UNIFORM, AQ_FINE_MASKING, AQ_FINE_UNIFORM, AQ_EDGE_REFINE = range(6)

EFFORT_FAST = 0
EFFORT_BALANCED = 1
MODE_NAVIGATOR = 0
MODE_BISECTION = 1
MODE_PRICED = 2


class Quantizer(ctypes.Structure):
    """jpxlg_quantizer."""

    _fields_ = [
        ("global_scale", ctypes.c_uint32),
        ("hf_mul", ctypes.c_uint32),
        ("quant_lf", ctypes.c_uint32),
        ("x_qm_scale", ctypes.c_uint32),
        ("b_qm_scale", ctypes.c_uint32),
        ("truncate", ctypes.c_int32),
        ("lambda_scale", ctypes.c_float),
        ("zero_token_bits", ctypes.c_float),
    ]


class Restoration(ctypes.Structure):
    """jpxlg_restoration."""

    _fields_ = [
        ("epf_iters", ctypes.c_int32),
        ("gaborish", ctypes.c_int32),
        ("sharpness", ctypes.c_int32),
    ]


class FrameDesc(ctypes.Structure):
    """jpxlg_frame_desc: what the codestream signals besides the plan."""

    _fields_ = [(n, ctypes.c_uint32) for n in ("width", "height", "bits_per_sample", "global_scale", "quant_lf",
                                                 "x_qm_scale", "b_qm_scale")] + [
        ("x_factor_lf", ctypes.c_int32), ("b_factor_lf", ctypes.c_int32),
        ("gaborish", ctypes.c_uint32), ("epf_iters", ctypes.c_uint32), ("sharpness", ctypes.c_uint32)]


class EmitSizes(ctypes.Structure):
    """jpxlg_emit_sizes (CodestreamSizing)."""

    _fields_ = [(n, ctypes.c_uint64) for n in ("total", "image_headers", "frame_header_bits", "toc_bits",
                                                 "lf_global", "lf_groups", "hf_global", "pass_groups")]


class QualityRequest(ctypes.Structure):
    """jpxlg_quality_request."""

    _fields_ = [
        ("target", ctypes.c_double),
        ("effort", ctypes.c_int32),
        ("max_probes", ctypes.c_int32),
        ("mode", ctypes.c_int32),
    ]


class Run(ctypes.Structure):
    """jpxlg_run: one varblock channel's token-ready run."""

    _fields_ = [("start", ctypes.c_uint64), ("count", ctypes.c_uint32), ("nonzeros", ctypes.c_uint32)]


class Varblock(ctypes.Structure):
    """jpxlg_varblock: one entry of the QuantizedFrameIr varblock order."""

    _fields_ = [
        ("lf_group", ctypes.c_uint32),
        ("bx", ctypes.c_uint32),
        ("by", ctypes.c_uint32),
        ("side", ctypes.c_uint32),
        ("hf_mul", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("start", ctypes.c_uint64),
    ]


class SourceFeatures(ctypes.Structure):
    """jpxlg_source_features (SourceFeatures + PreanalysisFeatures)."""

    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("grayscale", ctypes.c_int32)] + [
        (n, ctypes.c_float) for n in ("luma_variance_q10", "luma_variance_q50", "luma_variance_q90",
                                      "chroma_variance_q50", "flat_fraction", "edge_proxy")] + [
        ("has_preanalysis", ctypes.c_int32), ("cells", ctypes.c_uint32), ("share", ctypes.c_float * 7)] + [
        (n, ctypes.c_float) for n in ("log2var_mean", "log2var_std", "log2var_q10", "log2var_q90", "edge_mean",
                                      "edge_q90", "flat_mean", "noise_mean", "noise_q90", "chroma_log2_mean",
                                      "chroma_log2_q90", "orient_entropy_mean", "dir_dominance_mean",
                                      "axis_aligned_mean", "hetero_frac")]


class TransformSummary(ctypes.Structure):
    """jpxlg_transform_summary (TransformFeatureSummary)."""

    _fields_ = [("blocks", ctypes.c_uint64)] + [
        (n, ctypes.c_double) for n in ("ln_ac_y_mean", "ln_ac_y_q50", "ln_ac_y_q90", "ln_ac_y_q99",
                                       "high_low_ratio", "directional_asymmetry", "chroma_ac_ratio",
                                       "near_zero_frac_1e3", "near_zero_frac_1e2", "dc_variance_y")]


class CrossingPrediction(ctypes.Structure):
    """jpxlg_crossing_prediction (QualityPredictionV2)."""

    _fields_ = [("median_rung", ctypes.c_uint32), ("candidate_rung", ctypes.c_uint32),
                ("interval_low", ctypes.c_uint32), ("interval_high", ctypes.c_uint32),
                ("local_loss_exponent", ctypes.c_double), ("saturation_risk", ctypes.c_double),
                ("ood_flags", ctypes.c_uint64), ("fallback", ctypes.c_int32),
                ("large_frame_only_ood", ctypes.c_int32), ("model", ctypes.c_int32)]


OOD_TINY_FRAME = 1 << 62
OOD_CASE_SPREAD = 1 << 63
FALLBACK_NONE, FALLBACK_OOD_FEATURE, FALLBACK_WIDE_INTERVAL, FALLBACK_SATURATION_RISK = 0, 1, 2, 3
PREDICTOR_DISPATCH, PREDICTOR_CASES, PREDICTOR_KNOTS = 0, 1, 2


class RateRequest(ctypes.Structure):
    """jpxlg_rate_request."""

    _fields_ = [("target_bytes", ctypes.c_uint64), ("effort", ctypes.c_int32), ("max_prices", ctypes.c_uint32)]


class RateResult(ctypes.Structure):
    """jpxlg_rate_result."""

    _fields_ = [("rung", ctypes.c_uint32), ("global_scale", ctypes.c_uint32), ("hf_mul", ctypes.c_uint32),
                ("quant_lf", ctypes.c_uint32), ("bytes", ctypes.c_uint64), ("saturated", ctypes.c_int32),
                ("prices", ctypes.c_int32), ("trace_rungs", ctypes.c_uint32 * 64),
                ("trace_bytes", ctypes.c_uint64 * 64), ("floor_bytes", ctypes.c_uint64)]


class QualityResult(ctypes.Structure):
    """jpxlg_quality_result."""

    _fields_ = [
        ("score", ctypes.c_double),
        ("global_scale", ctypes.c_uint32),
        ("hf_mul", ctypes.c_uint32),
        ("quant_lf", ctypes.c_uint32),
        ("rung", ctypes.c_uint32),
        ("met", ctypes.c_int32),
        ("probes", ctypes.c_int32),
        ("probe_scores", ctypes.c_double * 16),
        ("probe_rungs", ctypes.c_uint32 * 16),
        ("probe_ms", ctypes.c_double * 16),
        ("status", ctypes.c_int32),
        ("predicted_rung", ctypes.c_uint32),
        ("structural_builds", ctypes.c_int32),
        ("structure_rebuilt", ctypes.c_int32),
        ("probe_fresh", ctypes.c_uint8 * 16),
        ("exact_prices", ctypes.c_int32),
        ("bytes", ctypes.c_uint64),
    ]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_u32 = ctypes.c_uint32
_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes)
SIGNATURES = {
    "jpxlg_ctx_create": (_i32, [_i32, ctypes.POINTER(_vp)]),
    "jpxlg_ctx_destroy": (None, [_vp]),
    "jpxlg_last_error": (ctypes.c_char_p, [_vp]),
    "jpxlg_image_footprint": (ctypes.c_uint64, [_u32, _u32]),
    "jpxlg_image_upload_srgb": (_i32, [_vp, _u32, _u32, _u32, _vp, ctypes.POINTER(_vp)]),
    "jpxlg_image_from_device": (_i32, [_vp, _u32, _u32, _u32, _vp, ctypes.POINTER(_vp)]),
    "jpxlg_image_destroy": (None, [_vp]),
    "jpxlg_image_info": (_i32, [_vp, _u32p, _u32p, _i32p, _u32p]),
    "jpxlg_set_structure": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "jpxlg_plan_structure": (_i32, [_vp, ctypes.POINTER(Quantizer), _i32, _i32]),
    "jpxlg_quantize": (_i32, [_vp, ctypes.POINTER(Quantizer)]),
    "jpxlg_set_dct_form": (_i32, [_vp, _i32]),
    "jpxlg_precondition_gaborish": (_i32, [_vp]),
    "jpxlg_render_score": (_i32, [_vp, ctypes.POINTER(Restoration), _dp, _dp, _dp, _u32p]),
    "jpxlg_plan_download": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jpxlg_image_download": (_i32, [_vp, _i32, _vp, ctypes.c_size_t]),
    "jpxlg_plan_download_ir": (_i32, [_vp, _vp, _vp, _vp, _u32p]),
    "jpxlg_emit_codestream": (_i32, [ctypes.POINTER(FrameDesc), _vp, _vp, _vp, _u32, _vp, _vp, _i32, _vp,
                                     ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(EmitSizes)]),
    "jpxlg_plan_download_runs": (_i32, [_vp, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _vp, _vp, _vp,
                                        _u32p]),
    "jpxlg_emit_codestream_runs": (_i32, [ctypes.POINTER(FrameDesc), _vp, ctypes.c_size_t, _vp, _vp, _vp, _u32, _vp,
                                          _vp, _i32, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                          ctypes.POINTER(EmitSizes)]),
    "jpxlg_encode_bytes": (_i32, [_vp, ctypes.POINTER(RateRequest), ctypes.POINTER(RateResult)]),
    "jpxlg_plan_frame_desc": (_i32, [_vp, ctypes.POINTER(Restoration), ctypes.POINTER(FrameDesc)]),
    "jpxlg_encode_emit": (_i32, [_vp, ctypes.POINTER(Restoration), _i32, _vp, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(EmitSizes)]),
    "jpxlg_set_plan": (_i32, [_vp, ctypes.POINTER(Quantizer), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jpxlg_encode_quality": (_i32, [_vp, ctypes.POINTER(QualityRequest), ctypes.POINTER(QualityResult)]),
    "jpxlg_rung_quantizer": (_i32, [_u32, _i32, ctypes.POINTER(Quantizer)]),
    "jpxlg_batch_probe": (_i32, [ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.POINTER(Quantizer),
                                 ctypes.POINTER(Restoration), ctypes.POINTER(ctypes.c_double)]),
    "jpxlg_batch_encode_quality": (_i32, [ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.POINTER(QualityRequest),
                                          ctypes.POINTER(QualityResult)]),
    "jpxlg_server_create": (_i32, [_vp, ctypes.POINTER(_vp)]),
    "jpxlg_server_submit": (_i32, [_vp, _vp, ctypes.POINTER(QualityRequest), ctypes.POINTER(ctypes.c_uint64)]),
    "jpxlg_server_submit_srgb": (_i32, [_vp, _u32, _u32, _u32, _vp, ctypes.POINTER(QualityRequest),
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "jpxlg_server_set_budget": (_i32, [_vp, ctypes.c_uint64]),
    "jpxlg_server_completed": (_i32, [_vp, ctypes.POINTER(ctypes.c_uint64), _u32, _u32p]),
    "jpxlg_server_admission": (_i32, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), _u32p]),
    "jpxlg_server_image": (_i32, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "jpxlg_server_step": (_i32, [_vp]),
    "jpxlg_image_quality_features": (_i32, [_vp, ctypes.POINTER(SourceFeatures), ctypes.POINTER(TransformSummary)]),
    "jpxlg_predict_crossing": (_i32, [ctypes.POINTER(SourceFeatures), ctypes.POINTER(TransformSummary), ctypes.c_double, ctypes.c_double,
                                      _i32, ctypes.POINTER(CrossingPrediction)]),
    "jpxlg_server_poll": (_i32, [_vp, ctypes.c_uint64, ctypes.POINTER(QualityResult)]),
    "jpxlg_server_release": (_i32, [_vp, ctypes.c_uint64]),
    "jpxlg_server_pending": (_i32, [_vp]),
    "jpxlg_server_destroy": (None, [_vp]),
    "jpxlg_score_pair": (_i32, [_vp, _u32, _u32, _vp, _vp, _dp, _dp, _u32p]),
    "jpxlg_score_pair_noalloc": (_i32, [_vp, _u32, _u32, _vp, _vp, _dp, _dp, _u32p]),
    "jpxlg_dct_blocks": (_i32, [_vp, _i32, _i32, _u32, _vp, _vp]),
    "jpxlg_epf_step": (_i32, [_vp, _i32, _u32, _u32, _vp, _vp, _vp]),
    "jpxlg_gaborish": (_i32, [_vp, _u32, _u32, _vp, _vp]),
    "jpxlg_set_aq": (_i32, [_vp, _i32, ctypes.POINTER(ctypes.c_int32)]),
    "jpxlg_plan_aq": (_i32, [_vp, _vp, ctypes.POINTER(Quantizer), ctypes.POINTER(ctypes.c_int32)]),
    "jpxlg_graph_create": (_i32, [ctypes.c_uint32, _vp, ctypes.POINTER(Restoration), ctypes.POINTER(_vp)]),
    "jpxlg_graph_probe": (_i32, [_vp, ctypes.POINTER(Quantizer), _dp]),
    "jpxlg_graph_destroy": (None, [_vp]),
    "jpxlg_autotune": (_i32, [_vp, ctypes.c_uint32, ctypes.c_uint32, _i32p, _i32p]),
    "jpxlg_profile_enable": (_i32, [_vp, _i32]),
    "jpxlg_profile_only": (_i32, [_vp, ctypes.c_char_p]),
    "jpxlg_profile_read": (_i32, [_vp, _vp, _dp, ctypes.POINTER(ctypes.c_uint64), _i32]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Loads libjpxlg.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("JPXLG_LIBRARY") or LIB_PATH  # an alternative build of the same ABI
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class JpxlgError(RuntimeError):
    """A non-zero status from the C ABI, with the context's last error."""

    def __init__(self, code: int, message: str):
        super().__init__(f"jpxlg status {code}: {message}")
        self.code = code
