"""jpxl_amd: the MI355X (gfx950) VarDCT pixel path of JPXL and its SSIMULACRA2
quality controller, behind the C ABI of ``include/jpxlg.h``.

The Python layer mirrors the reference's call surface for this path:

* ``Encoder().with_quality(85).with_effort("balanced").encode_rgb8_reported(w, h, rgb)``
  (jpxl/src/lib.rs:507-794) -> the quantized plan the CPU entropy coder
  consumes, plus a report (the reference's ``EncodeReport::Perceptual``);
* ``score_pair(reference, candidate)`` (jpxl-perceptual/src/lib.rs:195);
* ``Image`` exposes the resident per-image state: ``plan_structure``,
  ``quantize``, ``render_score`` (the ``PerceptualEvaluator`` seam of
  jpxl-encode-policy/src/quality.rs:87-157), ``plan_download``.

Every call runs on the GPU through ``libjpxlg.so``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import JpxlgError, Quantizer, Restoration

__all__ = [
    "Context",
    "Image",
    "Encoder",
    "EncodeReport",
    "Plan",
    "Quantizer",
    "Restoration",
    "JpxlgError",
    "score_pair",
    "quality_quantizer",
    "batch_probe",
    "batch_encode_quality",
    "Graph",
    "Server",
]


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """One device + one HIP stream (jpxlg_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.jpxlg_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise JpxlgError(rc, "context creation failed (no usable GPU?)")
        self.handle = h
        # Images hold device memory and the context's stream; the library
        # requires every image to be destroyed before its context.
        self._images = weakref.WeakSet()

    def check(self, rc: int) -> None:
        if rc != 0:
            msg = self.lib.jpxlg_last_error(self.handle)
            raise JpxlgError(rc, msg.decode() if msg else "")

    def close(self) -> None:
        if self.handle:
            for img in list(self._images):
                img.close()
            self.lib.jpxlg_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, width: int, height: int):
        """Pick the fastest streaming-pass launch shape for chains of
        width x height pixels; returns (units per workgroup, swizzle)."""
        upw, sw = ctypes.c_int32(), ctypes.c_int32()
        self.check(self.lib.jpxlg_autotune(self.handle, width, height, ctypes.byref(upw), ctypes.byref(sw)))
        return upw.value, sw.value

    # --- per-kernel timing ---------------------------------------------
    def profile(self, on: bool = True) -> None:
        self.check(self.lib.jpxlg_profile_enable(self.handle, int(on)))

    def profile_only(self, name: str | None) -> None:
        """Time only launches of kernel `name` (None: every launch)."""
        self.check(self.lib.jpxlg_profile_only(self.handle, name.encode() if name else None))

    def profile_read(self) -> dict:
        """{kernel: (total_ms, launches)} since the last read (HIP events)."""
        cap = 64
        names = ctypes.create_string_buffer(48 * cap)
        ms = np.zeros(cap, np.float64)
        n_l = np.zeros(cap, np.uint64)
        n = self.lib.jpxlg_profile_read(self.handle, names, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        n_l.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), cap)
        if n < 0:
            self.check(n)
        out = {}
        raw = names.raw
        for i in range(n):
            nm = raw[48 * i : 48 * i + 48].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(n_l[i]))
        return out

    # --- stand-alone kernels -------------------------------------------
    def dct_blocks(self, blocks: np.ndarray, inverse: bool = False) -> np.ndarray:
        blocks = np.ascontiguousarray(blocks, dtype=np.float32)
        n, s, s2 = blocks.shape
        assert s == s2
        out = np.empty_like(blocks)
        self.check(self.lib.jpxlg_dct_blocks(self.handle, s, int(inverse), n, _ptr(blocks), _ptr(out)))
        return out

    def epf_step(self, step: int, planes: np.ndarray, sigma: np.ndarray) -> np.ndarray:
        planes = np.ascontiguousarray(planes, dtype=np.float32)
        sigma = np.ascontiguousarray(sigma, dtype=np.float32)
        _, h, w = planes.shape
        assert sigma.shape == ((h + 7) // 8, (w + 7) // 8)
        out = np.empty_like(planes)
        self.check(self.lib.jpxlg_epf_step(self.handle, step, w, h, _ptr(planes), _ptr(sigma), _ptr(out)))
        return out

    def gaborish(self, planes: np.ndarray) -> np.ndarray:
        planes = np.ascontiguousarray(planes, dtype=np.float32)
        _, h, w = planes.shape
        out = np.empty_like(planes)
        self.check(self.lib.jpxlg_gaborish(self.handle, w, h, _ptr(planes), _ptr(out)))
        return out

    def score_pair(self, reference: np.ndarray, candidate: np.ndarray):
        """SSIMULACRA2 of two planar linear-sRGB images (3, H, W)."""
        reference = np.ascontiguousarray(reference, dtype=np.float32)
        candidate = np.ascontiguousarray(candidate, dtype=np.float32)
        _, h, w = reference.shape
        score = ctypes.c_double()
        terms = np.zeros(108, np.float64)
        nterms = ctypes.c_uint32()
        self.check(self.lib.jpxlg_score_pair(self.handle, w, h, _ptr(reference), _ptr(candidate),
                                             ctypes.byref(score), terms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             ctypes.byref(nterms)))
        return score.value, terms[: nterms.value]


_DEFAULT_CTX: Context | None = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


def score_pair(reference: np.ndarray, candidate: np.ndarray) -> float:
    """jpxl_perceptual::score_pair on the GPU; planar linear sRGB (3, H, W)."""
    return default_context().score_pair(reference, candidate)[0]


def quality_quantizer(rung_fields: tuple[int, int], effort: str = "balanced", quant_lf: int = 4) -> Quantizer:
    """The quantizer a quality probe uses at (global_scale, hf_mul)."""
    gs, mul = rung_fields
    fast = effort == "fast"
    q = Quantizer()
    q.global_scale, q.hf_mul, q.quant_lf = gs, mul, quant_lf
    q.x_qm_scale = q.b_qm_scale = 2 if fast else 3
    q.truncate = 0 if fast else 1
    q.lambda_scale = 4.0
    q.zero_token_bits = 1.0
    return q


def rung_quantizer(rung: int, effort: str = "balanced") -> Quantizer:
    """The quantizer a quality search of `effort` probes at ladder rung `rung`
    (QuantizerChoice::at, quantizer_ladder.rs:202-260)."""
    q = Quantizer()
    lib = _lib.load()
    rc = lib.jpxlg_rung_quantizer(rung, _lib.EFFORT_FAST if effort == "fast" else _lib.EFFORT_BALANCED,
                                  ctypes.byref(q))
    if rc:
        raise JpxlgError(rc, "rung outside the ladder")
    return q


@dataclass
class Plan:
    """The quantized plan handed to the CPU entropy coder (QuantizedFrameIr)."""

    quant: np.ndarray        # (3, 8*bh, 8*bw) int32, footprint layout
    lf: np.ndarray           # (3, bh, bw) int32
    block_map: np.ndarray    # (bh, bw) uint8: 8/16/32 at varblock origins, 0 covered
    cfl_x: np.ndarray        # (th, tw) int8
    cfl_b: np.ndarray        # (th, tw) int8
    lf_factors: tuple        # (x_factor_lf - 128, b_factor_lf - 128)


class Image:
    """One image resident in HBM: XYB planes, coefficients, metric reference."""

    def __init__(self, ctx: Context, rgb: np.ndarray, bits: int | None = None, device_ptr: int | None = None,
                 shape: tuple | None = None):
        self.handle = None
        self.ctx = ctx
        lib = ctx.lib
        h = ctypes.c_void_p()
        if device_ptr is not None:
            height, width = shape
            self.bits = bits
            rc = lib.jpxlg_image_from_device(ctx.handle, width, height, bits, ctypes.c_void_p(device_ptr),
                                             ctypes.byref(h))
        else:
            rgb = np.ascontiguousarray(rgb)
            height, width, three = rgb.shape
            assert three == 3
            if bits is None:
                bits = 8 if rgb.dtype == np.uint8 else 16
            self.bits = bits
            want = np.uint8 if bits == 8 else np.uint16
            rgb = rgb.astype(want, copy=False)
            rc = lib.jpxlg_image_upload_srgb(ctx.handle, width, height, bits, _ptr(rgb), ctypes.byref(h))
        ctx.check(rc)
        self.handle = h
        ctx._images.add(self)
        self.width, self.height = width, height
        bx, by, gray, scales = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_uint32()
        ctx.check(lib.jpxlg_image_info(h, ctypes.byref(bx), ctypes.byref(by), ctypes.byref(gray),
                                       ctypes.byref(scales)))
        self.bw, self.bh, self.grayscale, self.scales = bx.value, by.value, bool(gray.value), scales.value
        self.tw, self.th = (self.bw + 7) // 8, (self.bh + 7) // 8

    @classmethod
    def _borrowed(cls, ctx: Context, handle) -> "Image":
        """A view of an image another owner keeps (closing it frees nothing)."""
        self = cls.__new__(cls)
        self.handle = None
        self.ctx = ctx
        self._owner = False
        bx, by, gray, scales = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_uint32()
        ctx.check(ctx.lib.jpxlg_image_info(handle, ctypes.byref(bx), ctypes.byref(by), ctypes.byref(gray),
                                           ctypes.byref(scales)))
        self.handle = handle
        self.bw, self.bh, self.grayscale, self.scales = bx.value, by.value, bool(gray.value), scales.value
        self.tw, self.th = (self.bw + 7) // 8, (self.bh + 7) // 8
        self.width, self.height, self.bits = 8 * self.bw, 8 * self.bh, None
        return self

    def close(self):
        if self.handle:
            if self.ctx.handle and getattr(self, "_owner", True):
                self.ctx.lib.jpxlg_image_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_structure(self, block_map: np.ndarray, cfl_x=None, cfl_b=None, lf_factors=(0, 0)):
        bm = np.ascontiguousarray(block_map, dtype=np.uint8)
        assert bm.shape == (self.bh, self.bw)
        cx = None if cfl_x is None else np.ascontiguousarray(cfl_x, dtype=np.int8)
        cb = None if cfl_b is None else np.ascontiguousarray(cfl_b, dtype=np.int8)
        lf = np.asarray(lf_factors, dtype=np.int32)
        self.ctx.check(self.ctx.lib.jpxlg_set_structure(
            self.handle, _ptr(bm), None if cx is None else _ptr(cx), None if cb is None else _ptr(cb), _ptr(lf)))

    def plan_structure(self, q: Quantizer, cover_mode: int = _lib.COVER_FIXED_DCT8, enable_cfl: bool = True):
        self.ctx.check(self.ctx.lib.jpxlg_plan_structure(self.handle, ctypes.byref(q), cover_mode, int(enable_cfl)))

    def quantize(self, q: Quantizer):
        self.ctx.check(self.ctx.lib.jpxlg_quantize(self.handle, ctypes.byref(q)))

    def set_aq(self, mode: int) -> bool:
        """Build the AQ field of `mode` (_lib.AQ_*) from the source planes;
        returns whether it is active (not off, not neutral).  Drops the
        structure; call before precondition_gaborish."""
        active = ctypes.c_int32()
        self.ctx.check(self.ctx.lib.jpxlg_set_aq(self.handle, int(mode), ctypes.byref(active)))
        return bool(active.value)

    def plan_aq(self):
        """-> (per-block HfMul (bh, bw) uint32, wire Quantizer, refine_only) of
        the last quantized plan."""
        muls = np.zeros((self.bh, self.bw), np.uint32)
        wire = Quantizer()
        ro = ctypes.c_int32()
        self.ctx.check(self.ctx.lib.jpxlg_plan_aq(self.handle, _ptr(muls), ctypes.byref(wire), ctypes.byref(ro)))
        return muls, wire, bool(ro.value)

    def precondition_gaborish(self):
        """Replace the source XYB by its Gaborish inverse (drops the structure)."""
        self.ctx.check(self.ctx.lib.jpxlg_precondition_gaborish(self.handle))

    def set_dct_form(self, form: int):
        """_lib.DCT_EXACT (the reference's butterflies), DCT_FP32 or DCT_BF16X3
        (matrix-core contractions of the precision study)."""
        self.ctx.check(self.ctx.lib.jpxlg_set_dct_form(self.handle, int(form)))

    def render_score(self, restoration: Restoration | None = None):
        r = restoration or Restoration(1, 0, 7)
        score, raw = ctypes.c_double(), ctypes.c_double()
        terms = np.zeros(108, np.float64)
        n = ctypes.c_uint32()
        self.ctx.check(self.ctx.lib.jpxlg_render_score(
            self.handle, ctypes.byref(r), ctypes.byref(score), ctypes.byref(raw),
            terms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n)))
        return score.value, raw.value, terms[: n.value]

    def quality_features(self):
        """(SourceFeatures, TransformSummary) the quality search's first
        guess reads (ctypes structs)."""
        sf, ts = _lib.SourceFeatures(), _lib.TransformSummary()
        self.ctx.check(self.ctx.lib.jpxlg_image_quality_features(self.handle, ctypes.byref(sf), ctypes.byref(ts)))
        return sf, ts

    def plan_download(self) -> Plan:
        q = np.empty((3, 8 * self.bh, 8 * self.bw), np.int32)
        lf = np.empty((3, self.bh, self.bw), np.int32)
        bm = np.empty((self.bh, self.bw), np.uint8)
        cx = np.empty((self.th, self.tw), np.int8)
        cb = np.empty((self.th, self.tw), np.int8)
        lff = np.empty(2, np.int32)
        self.ctx.check(self.ctx.lib.jpxlg_plan_download(self.handle, _ptr(q), _ptr(lf), _ptr(bm), _ptr(cx),
                                                        _ptr(cb), _ptr(lff)))
        return Plan(q, lf, bm, cx, cb, (int(lff[0]), int(lff[1])))

    def encode_emit(self, restoration: "Restoration | None" = None, threads: int = 0):
        """The JPEG XL codestream of the last quantized plan (attach_entropy +
        emit on the host): (bytes, EmitSizes).  ``restoration`` defaults to
        that of the image's last render."""
        lib = self.ctx.lib
        n = ctypes.c_size_t()
        sizes = _lib.EmitSizes()
        rp = ctypes.byref(restoration) if restoration is not None else None
        self.ctx.check(lib.jpxlg_encode_emit(self.handle, rp, threads, None, 0, ctypes.byref(n), ctypes.byref(sizes)))
        out = np.empty(n.value, np.uint8)
        self.ctx.check(lib.jpxlg_encode_emit(self.handle, rp, threads, _ptr(out), out.size, ctypes.byref(n),
                                             ctypes.byref(sizes)))
        return out.tobytes(), sizes

    def plan_download_ir(self):
        """The QuantizedFrameIr hand-off: (coefficient arena, per-group LF
        planes, varblock table as a list of dicts in the IR's order)."""
        arena = np.empty(192 * self.bw * self.bh, np.int32)
        lfg = np.empty(3 * self.bw * self.bh, np.int32)
        table = (_lib.Varblock * (self.bw * self.bh))()
        n = ctypes.c_uint32()
        self.ctx.check(self.ctx.lib.jpxlg_plan_download_ir(self.handle, _ptr(arena), _ptr(lfg), table,
                                                           ctypes.byref(n)))
        vbs = [dict(lf_group=v.lf_group, bx=v.bx, by=v.by, side=v.side, hf_mul=v.hf_mul, start=v.start)
               for v in table[: n.value]]
        return arena, lfg, vbs

    def plan_download_runs(self):
        """The token-ready hand-off: (run arena, runs (3 per varblock, ctypes
        array of Run), per-group LF planes, varblock table as in
        plan_download_ir)."""
        bpx = self.bw * self.bh
        arena = np.empty(189 * bpx, np.int32)
        runs = (_lib.Run * (3 * bpx))()
        lfg = np.empty(3 * bpx, np.int32)
        table = (_lib.Varblock * bpx)()
        n, used = ctypes.c_uint32(), ctypes.c_size_t()
        self.ctx.check(self.ctx.lib.jpxlg_plan_download_runs(self.handle, _ptr(arena), arena.size, ctypes.byref(used),
                                                             runs, _ptr(lfg), table, ctypes.byref(n)))
        vbs = [dict(lf_group=v.lf_group, bx=v.bx, by=v.by, side=v.side, hf_mul=v.hf_mul, start=v.start)
               for v in table[: n.value]]
        return arena[: used.value].copy(), runs[: 3 * n.value], lfg, vbs

    def set_plan(self, wire: Quantizer, plan: "Plan", hf_mul: np.ndarray | None = None) -> None:
        """PerceptualEvaluator::evaluate's input: a plan built elsewhere becomes
        this image's structure and levels (the next render_score renders it)."""
        bm = np.ascontiguousarray(plan.block_map, np.uint8)
        cx = np.ascontiguousarray(plan.cfl_x, np.int8)
        cb = np.ascontiguousarray(plan.cfl_b, np.int8)
        lff = np.array(plan.lf_factors, np.int32)
        q = np.ascontiguousarray(plan.quant, np.int32)
        lf = np.ascontiguousarray(plan.lf, np.int32)
        muls = None if hf_mul is None else np.ascontiguousarray(hf_mul, np.uint32)
        self.ctx.check(self.ctx.lib.jpxlg_set_plan(self.handle, ctypes.byref(wire), _ptr(bm), _ptr(cx), _ptr(cb),
                                                   _ptr(lff), _ptr(muls) if muls is not None else None, _ptr(q),
                                                   _ptr(lf)))

    def plan_download_map(self) -> np.ndarray:
        """The current block map (valid right after plan_structure)."""
        bm = np.empty((self.bh, self.bw), np.uint8)
        q = self.ctx.lib
        self.ctx.check(q.jpxlg_plan_download(self.handle, None, None, _ptr(bm), None, None, None))
        return bm

    def download(self, which: int) -> np.ndarray:
        if which in (_lib.BUF_XYB, _lib.BUF_RENDER_XYB, _lib.BUF_LINEAR, _lib.BUF_REF_XYB0):
            out = np.empty((3, self.height, self.width), np.float32)
        elif which == _lib.BUF_COEFF:
            out = np.empty((3, 8 * self.bh, 8 * self.bw), np.float32)
        elif which in (_lib.BUF_SIGMA, _lib.BUF_AQ_FIELD):
            out = np.empty((self.bh, self.bw), np.float32)
        else:
            raise ValueError(which)
        self.ctx.check(self.ctx.lib.jpxlg_image_download(self.handle, which, _ptr(out), out.nbytes))
        return out

    def encode_bytes(self, target_bytes: int, effort: str = "balanced", max_prices: int = 0):
        """rate::search_frame over a byte target: the RateResult; the chosen
        plan stays quantized (jpxlg_encode_emit writes it)."""
        return _encode_bytes(self, target_bytes, effort, max_prices)

    def encode_quality(self, target: float, effort: str = "balanced", max_probes: int = 0, mode: int = 0):
        req = _lib.QualityRequest(target, _lib.EFFORT_FAST if effort == "fast" else _lib.EFFORT_BALANCED,
                                  max_probes, mode)
        res = _lib.QualityResult()
        self.ctx.check(self.ctx.lib.jpxlg_encode_quality(self.handle, ctypes.byref(req), ctypes.byref(res)))
        return res


def _encode_bytes(img, target_bytes: int, effort: str = "balanced", max_prices: int = 0):
    req = _lib.RateRequest(int(target_bytes), _lib.EFFORT_FAST if effort == "fast" else _lib.EFFORT_BALANCED,
                           max_prices)
    res = _lib.RateResult()
    img.ctx.check(img.ctx.lib.jpxlg_encode_bytes(img.handle, ctypes.byref(req), ctypes.byref(res)))
    return res


def predict_crossing(source, transform, target: float, reserve: float, which: int = 0):
    """quality_prediction::predict (which 0), predict_cases (1) or
    predict_v2 (2): a CrossingPrediction, or None.  Host-only."""
    out = _lib.CrossingPrediction()
    rc = _lib.load().jpxlg_predict_crossing(ctypes.byref(source), ctypes.byref(transform), float(target),
                                            float(reserve), int(which), ctypes.byref(out))
    if rc < 0:
        raise JpxlgError(rc, "jpxlg_predict_crossing")
    return out if rc == 1 else None


def _image_array(images):
    if not images:
        raise ValueError("empty batch")
    ctx = images[0].ctx
    if any(im.ctx is not ctx for im in images):
        raise ValueError("a batch's images must share one Context")
    arr = (ctypes.c_void_p * len(images))(*[im.handle.value for im in images])
    return ctx, arr


def batch_probe(images, quantizers, restoration: Restoration | None = None) -> np.ndarray:
    """One probe of every image (each with its own Quantizer), one launch per
    stage over the batch; returns the SSIMULACRA2 scores."""
    ctx, arr = _image_array(images)
    if len(quantizers) != len(images):
        raise ValueError("one quantizer per image")
    qs = (Quantizer * len(images))(*quantizers)
    r = restoration or Restoration(1, 0, 7)
    scores = np.zeros(len(images), np.float64)
    ctx.check(ctx.lib.jpxlg_batch_probe(len(images), arr, qs, ctypes.byref(r),
                                        scores.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return scores


class Graph:
    """The probe chain of a fixed set of images captured as a HIP graph;
    ``probe(quantizers)`` replays it (one quantizer per image) and returns the
    scores.  Valid while the images keep their structure."""

    def __init__(self, images, restoration: Restoration | None = None):
        self.ctx, arr = _image_array(images)
        self.images = list(images)
        self.handle = ctypes.c_void_p()
        r = restoration or Restoration(1, 0, 7)
        self.ctx.check(self.ctx.lib.jpxlg_graph_create(len(self.images), arr, ctypes.byref(r),
                                                       ctypes.byref(self.handle)))

    def probe(self, quantizers) -> np.ndarray:
        if len(quantizers) != len(self.images):
            raise ValueError("one quantizer per image")
        qs = (Quantizer * len(self.images))(*quantizers)
        scores = np.zeros(len(self.images), np.float64)
        self.ctx.check(self.ctx.lib.jpxlg_graph_probe(self.handle, qs,
                                                      scores.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return scores

    def close(self):
        if self.handle:
            self.ctx.lib.jpxlg_graph_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_encode_quality(images, targets, effort: str = "balanced", max_probes: int = 0, mode: int = 0):
    """The quality search of every image at once (each its own target): each
    round of probes of the images still searching is one batched chain.
    Returns one QualityResult per image; each chosen plan stays quantized."""
    ctx, arr = _image_array(images)
    eff = _lib.EFFORT_FAST if effort == "fast" else _lib.EFFORT_BALANCED
    reqs = (_lib.QualityRequest * len(images))(*[_lib.QualityRequest(float(t), eff, max_probes, mode)
                                                  for t in targets])
    res = (_lib.QualityResult * len(images))()
    ctx.check(ctx.lib.jpxlg_batch_encode_quality(len(images), arr, reqs, res))
    return list(res)


class Server:
    """Continuous batching of quality requests (jpxlg_server): submit images
    with their targets at any time; each step() runs one round of probes of
    every active request as one batched chain."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._held = {}  # ticket -> host samples of a request given as samples
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.jpxlg_server_create(ctx.handle, ctypes.byref(h)))
        self.handle = h

    def submit(self, image: "Image", target: float, effort: str = "balanced", max_probes: int = 0,
               mode: int = 0) -> int:
        req = _lib.QualityRequest(float(target), _lib.EFFORT_FAST if effort == "fast" else _lib.EFFORT_BALANCED,
                                  max_probes, mode)
        t = ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.jpxlg_server_submit(self.handle, image.handle, ctypes.byref(req), ctypes.byref(t)))
        return t.value

    def submit_srgb(self, rgb: np.ndarray, target: float, effort: str = "balanced", bits: int | None = None,
                    max_probes: int = 0, mode: int = 0) -> int:
        """A request as host samples: the server admits it (creating its image)
        when its projected footprint fits the budget; the samples are held
        until the request is released."""
        rgb = np.ascontiguousarray(rgb)
        height, width, three = rgb.shape
        assert three == 3
        if bits is None:
            bits = 8 if rgb.dtype == np.uint8 else 16
        rgb = rgb.astype(np.uint8 if bits == 8 else np.uint16, copy=False)
        req = _lib.QualityRequest(float(target), _lib.EFFORT_FAST if effort == "fast" else _lib.EFFORT_BALANCED,
                                  max_probes, mode)
        t = ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.jpxlg_server_submit_srgb(self.handle, width, height, bits, _ptr(rgb),
                                                             ctypes.byref(req), ctypes.byref(t)))
        self._held[t.value] = rgb
        return t.value

    def completed(self, cap: int = 4096) -> list[int]:
        """Tickets of the requests completed since the last call."""
        buf = (ctypes.c_uint64 * cap)()
        n = ctypes.c_uint32()
        self.ctx.check(self.ctx.lib.jpxlg_server_completed(self.handle, buf, cap, ctypes.byref(n)))
        return list(buf[: n.value])

    def set_budget(self, nbytes: int) -> None:
        self.ctx.check(self.ctx.lib.jpxlg_server_set_budget(self.handle, int(nbytes)))

    def admission(self) -> tuple[int, int, int]:
        """(projected footprint of the admitted requests, budget, queued requests)."""
        used, budget, queued = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
        self.ctx.check(self.ctx.lib.jpxlg_server_admission(self.handle, ctypes.byref(used), ctypes.byref(budget),
                                                           ctypes.byref(queued)))
        return used.value, budget.value, queued.value

    def plan(self, ticket: int) -> "Plan":
        """The chosen plan of a completed request (its image's quantized planes)."""
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.jpxlg_server_image(self.handle, ticket, ctypes.byref(h)))
        return Image._borrowed(self.ctx, h).plan_download()

    def step(self) -> int:
        rc = self.ctx.lib.jpxlg_server_step(self.handle)
        if rc < 0:
            self.ctx.check(rc)
        return rc

    def poll(self, ticket: int):
        """The QualityResult once the request is done, else None."""
        res = _lib.QualityResult()
        rc = self.ctx.lib.jpxlg_server_poll(self.handle, ticket, ctypes.byref(res))
        if rc < 0:
            self.ctx.check(rc)
        return res if rc == 1 else None

    def release(self, ticket: int) -> None:
        self.ctx.check(self.ctx.lib.jpxlg_server_release(self.handle, ticket))
        self._held.pop(ticket, None)

    def pending(self) -> int:
        return self.ctx.lib.jpxlg_server_pending(self.handle)

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.jpxlg_server_destroy(self.handle)
            self.handle = None
        self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class EncodeReport:
    """The perceptual half of the reference's EncodeReport (jpxl/src/lib.rs:353)."""

    target: float
    achieved: float
    met: bool
    global_scale: int
    hf_mul: int
    quant_lf: int
    probes: list = field(default_factory=list)
    bytes: int = 0


class Encoder:
    """Builder mirroring jpxl::Encoder for the quality / fixed-quantizer paths.

    ``encode_rgb8`` returns the JPEG XL codestream (``bytes``) and
    ``encode_rgb8_reported`` ``(bytes, EncodeReport)``, as the facade's
    ``Result<Vec<u8>>`` / ``Result<(Vec<u8>, EncodeReport)>``; ``plan_rgb8``
    returns the quantized plan itself (what the reference hands to its
    entropy coder).
    """

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.quality: float | None = None
        self.global_scale: int | None = None
        self.effort = "balanced"
        self.threads = 0
        self.target_bytes: int | None = None
        self.target_bpp: float | None = None

    def with_target_bytes(self, target_bytes: int) -> "Encoder":
        """Lossy VarDCT to an exact byte ceiling (jpxl/src/lib.rs:546)."""
        if int(target_bytes) <= 0:
            raise ValueError("target byte count must be greater than zero")
        self.target_bytes, self.target_bpp, self.quality, self.global_scale = int(target_bytes), None, None, None
        return self

    def with_target_bpp(self, bits_per_pixel: float) -> "Encoder":
        """Lossy VarDCT to a bits-per-pixel ceiling (jpxl/src/lib.rs:534)."""
        if not (bits_per_pixel > 0.0 and bits_per_pixel < float("inf")):
            raise ValueError("target bits per pixel must be finite and greater than zero")
        self.target_bpp, self.target_bytes, self.quality, self.global_scale = float(bits_per_pixel), None, None, None
        return self

    def with_quality(self, quality: float) -> "Encoder":
        if not (0.0 <= quality <= 100.0):
            raise ValueError("quality must be within 0..=100")
        self.quality, self.global_scale = float(quality), None
        self.target_bytes = self.target_bpp = None
        return self

    with_ssimulacra2_score = with_quality

    def with_global_scale(self, global_scale: int) -> "Encoder":
        if not (1 <= global_scale <= 73728):
            raise ValueError("global_scale outside 1..=73728")
        self.global_scale, self.quality = int(global_scale), None
        self.target_bytes = self.target_bpp = None
        return self

    def with_effort(self, effort: str) -> "Encoder":
        if effort not in ("fast", "balanced"):
            raise ValueError(f"unsupported effort {effort!r}")
        self.effort = effort
        return self

    def with_threads(self, threads: int) -> "Encoder":
        self.threads = int(threads)  # host workers of the codestream stage (0: every core)
        return self

    def _encode(self, rgb: np.ndarray, bits: int, emit: bool = True):
        img = Image(self.ctx, rgb, bits)
        try:
            if self.target_bytes is not None or self.target_bpp is not None:
                h, w = rgb.shape[:2]
                # RateTarget::bytes_for (request.rs:46-70)
                target = self.target_bytes if self.target_bytes is not None else \
                    int(math.floor(self.target_bpp * w * h / 8.0))
                res = img.encode_bytes(target, self.effort)
                report = EncodeReport(float("nan"), float("nan"), True, res.global_scale, res.hf_mul, res.quant_lf)
            elif self.quality is not None:
                # the reference's default build: finalists priced exactly
                res = img.encode_quality(self.quality, self.effort, mode=_lib.MODE_PRICED)
                report = EncodeReport(self.quality, res.score, bool(res.met), res.global_scale, res.hf_mul,
                                      res.quant_lf, [(res.probe_rungs[i], res.probe_scores[i]) for i in range(res.probes)])
            else:
                # EncodeRequest::for_fixed_quantizer over the defaults
                # (request.rs:855-880, :936-943): quant_lf 16, hf_mul 1, AQ
                # masking, hierarchical cover priced with the legacy rate
                # proxy and flat cell weights, CfL on, the nearest rule.
                q = Quantizer(self.global_scale or 32768, 1, 16, 2, 2, 0, 1.0, 1.0)
                img.set_aq(_lib.AQ_MASKING)
                img.plan_structure(q, _lib.COVER_HIERARCHICAL_LEGACY, True)
                img.quantize(q)
                _, wire, _ = img.plan_aq()
                report = EncodeReport(float("nan"), float("nan"), True, wire.global_scale, wire.hf_mul,
                                      wire.quant_lf)
                # the fixed-quantizer request's restoration (request.rs:905)
                img.render_score(Restoration(1, 0, 7))
            if emit:
                stream, sizes = img.encode_emit(None, self.threads)
                report.bytes = sizes.total
                return stream, report
            return img.plan_download(), report
        finally:
            img.close()

    def encode_rgb8_reported(self, width: int, height: int, rgb: bytes | np.ndarray):
        a = np.frombuffer(bytes(rgb), np.uint8) if not isinstance(rgb, np.ndarray) else rgb
        return self._encode(a.reshape(height, width, 3).astype(np.uint8), 8)

    def encode_rgb16_reported(self, width: int, height: int, rgb: np.ndarray, bits: int = 16):
        return self._encode(np.asarray(rgb, np.uint16).reshape(height, width, 3), bits)

    def plan_rgb8(self, width: int, height: int, rgb):
        """(Plan, EncodeReport): the quantized plan without the codestream."""
        a = np.frombuffer(bytes(rgb), np.uint8) if not isinstance(rgb, np.ndarray) else rgb
        return self._encode(a.reshape(height, width, 3).astype(np.uint8), 8, emit=False)

    def encode_rgb8(self, width: int, height: int, rgb):
        return self.encode_rgb8_reported(width, height, rgb)[0]

    def encode_rgb16(self, width: int, height: int, rgb, bits: int = 16):
        return self.encode_rgb16_reported(width, height, rgb, bits)[0]
