"""The whole quality encode on the CPU, as test infrastructure: the library's
navigator (jpxl_amd/csrc/controller.cpp compiled host-only, the same code the
GPU path runs) driving the oracle's probes (cover + CfL on fresh probes,
quantize, render, SSIMULACRA2), started where the library's crossing
predictor (or the legacy table) puts the first probe for the oracle's
features.  It is the CPU restatement of search_frame_perceptual
(jpxl-encode-policy/src/quality.rs:2405-2740) used as the bench's CPU
baseline and as the GPU navigator's end-to-end check; never shipped."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

import oracle_binding as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "host", "libcontroller_check.so")
CSRC = os.path.join(ROOT, "jpxl_amd", "csrc")
SRCS = [os.path.join(ROOT, "tests", "host", "controller_check.cpp"),
        os.path.join(CSRC, "controller.cpp"),
        os.path.join(CSRC, "metric_host.cpp"),
        os.path.join(CSRC, "predictor.cpp"),
        os.path.join(CSRC, "predictor_tables.S")]

PROBE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_double))


class Outcome(ctypes.Structure):
    _fields_ = [("rung", ctypes.c_uint32), ("score", ctypes.c_double), ("met", ctypes.c_int32),
                ("status", ctypes.c_int32), ("chosen_fresh", ctypes.c_int32), ("anchor_rung", ctypes.c_uint32),
                ("pixel_probes", ctypes.c_int32), ("structural_builds", ctypes.c_int32),
                ("priced", ctypes.c_int32), ("nprobes", ctypes.c_int32),
                ("probe_rungs", ctypes.c_uint32 * 32), ("probe_scores", ctypes.c_double * 32),
                ("probe_fresh", ctypes.c_int32 * 32)]


def build(force: bool = False) -> str:
    """Compiles the host-only controller shim (g++, a few seconds)."""
    deps = SRCS + [os.path.join(CSRC, "predictor_tables.bin")]
    if force or not os.path.exists(SO) or any(os.path.getmtime(s) > os.path.getmtime(SO) for s in deps):
        subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off", f"-Wa,-I{CSRC}",
                        "-o", SO] + SRCS, check=True)
    return SO


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        cc = ctypes.CDLL(build())
        d, u32 = ctypes.c_double, ctypes.c_uint32
        cc.cc_predicted_scale.restype = d
        cc.cc_predicted_scale.argtypes = [ctypes.c_float, ctypes.c_float, d]
        cc.cc_rung_for_scale.restype = u32
        cc.cc_rung_for_scale.argtypes = [d]
        cc.cc_features.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                   ctypes.POINTER(ctypes.c_float)]
        cc.cc_navigate_callback.restype = ctypes.c_int
        cc.cc_navigate_callback.argtypes = [d, ctypes.c_int, ctypes.c_int, u32, d, ctypes.c_int, PROBE_FN,
                                            ctypes.POINTER(Outcome)]
        from jpxl_amd import _lib as jl

        cc.cc_quality_start.argtypes = [ctypes.POINTER(jl.SourceFeatures), ctypes.POINTER(jl.TransformSummary), d,
                                        ctypes.c_int, u32, ctypes.POINTER(u32), ctypes.POINTER(d)]
        cc.cc_rung_policy.argtypes = [u32, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _LIB = cc
    return _LIB


def predicted_rung(xyb: np.ndarray, target: float) -> int:
    """quality.rs:2452-2460 from the oracle's per-atom luma variance."""
    cc = lib()
    var = np.ascontiguousarray(orc.atlas_variance(xyb)[1].reshape(-1), np.float32)
    q50, flat = ctypes.c_float(), ctypes.c_float()
    cc.cc_features(var.ctypes.data, var.size, ctypes.byref(q50), ctypes.byref(flat))
    return int(cc.cc_rung_for_scale(cc.cc_predicted_scale(q50.value, flat.value, target)))


def oracle_features(rgb: np.ndarray, bits: int, xyb: np.ndarray, gray: bool):
    """The first-guess features from the oracle (quality.rs:2449-2474), as
    the library's ctypes structs."""
    from jpxl_amd import _lib as jl

    sf = jl.SourceFeatures()
    for k, v in orc.source_features(xyb, gray).items():
        setattr(sf, k, v)
    pa = orc.preanalysis(rgb, bits)
    if pa is not None:
        sf.has_preanalysis = 1
        sf.cells = pa["cells"]
        for i in range(7):
            sf.share[i] = float(pa["share"][i])
        for k in orc.PREANALYSIS_FIELDS[1:]:
            setattr(sf, k, float(pa[k]))
    ts = jl.TransformSummary()
    t, _ = orc.transform_summary(xyb)
    for k, v in t.items():
        setattr(ts, k, v)
    return sf, ts


def quality_start(rgb: np.ndarray, bits: int, xyb: np.ndarray, gray: bool, target: float, effort: int):
    """(first rung, prior exponent or NaN) from the oracle's features."""
    sf, ts = oracle_features(rgb, bits, xyb, gray)
    r, b = ctypes.c_uint32(), ctypes.c_double()
    lib().cc_quality_start(ctypes.byref(sf), ctypes.byref(ts), target, effort, predicted_rung(xyb, target),
                           ctypes.byref(r), ctypes.byref(b))
    return int(r.value), float(b.value)


def rung_policy(rung: int, effort: int):
    cc = lib()
    f = np.zeros(6, np.uint32)
    t = np.zeros(3, np.int32)
    lz = np.zeros(2, np.float32)
    cc.cc_rung_policy(rung, effort, f.ctypes.data, t.ctypes.data, lz.ctypes.data)
    qp = orc.QuantParams(int(f[0]), int(f[1]), int(f[2]), int(f[3]), int(f[4]), int(f[5]), float(lz[0]),
                         float(lz[1]))
    epf, sharp = int(t[2]) // 16, int(t[2]) % 16
    return qp, int(t[0]), bool(t[1]), epf, sharp


def encode(rgb: np.ndarray, bits: int, target: float, effort: int = 1, mode: int = 0, max_probes: int = 0):
    """One whole quality encode on the CPU; returns (Outcome, final plan
    (quant, lf, bmap, cfl_x, cfl_b, lf_factors))."""
    h, w, _ = rgb.shape
    xyb, lin, gray = orc.prepare_frame(rgb, bits)
    pred, prior = quality_start(rgb, bits, xyb, gray, target, effort) if mode == 0 else (0, float("nan"))
    state = {}

    def structure(rung):
        qp, cover_mode, cfl, _, _ = rung_policy(rung, effort)
        bmap = orc.cover(xyb, qp) if cover_mode == 1 else orc.fixed_dct8_map(w, h)
        coeff = orc.forward_frame(xyb, bmap)
        if cfl and not gray:
            cx, cb, lff = orc.estimate_cfl(w, h, bmap, coeff, qp)
        else:
            th, tw = (bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8
            cx = cb = np.zeros((th, tw), np.int8)
            lff = (0, 0)
        return bmap, coeff, cx, cb, lff

    def plan_at(rung, fresh):
        if fresh or "anchor" not in state:
            st = structure(rung)
            if "anchor" not in state:
                state["anchor"] = st
            state["fresh", rung] = st
        st = state["fresh", rung] if fresh else state["anchor"]
        bmap, coeff, cx, cb, lff = st
        qp, _, _, epf, sharp = rung_policy(rung, effort)
        quant, lf = orc.quantize_frame(w, h, bmap, coeff, qp, cx, cb, lff)
        return (quant, lf, bmap, cx, cb, lff), qp, epf, sharp

    def probe(rung, fresh, score_out):
        try:
            (quant, lf, bmap, cx, cb, lff), qp, epf, sharp = plan_at(rung, bool(fresh))
            rp = orc.RenderParams(qp.global_scale, qp.hf_mul, qp.quant_lf, qp.x_qm, qp.b_qm, epf, 0, sharp, bits)
            olin, _ = orc.render(w, h, bmap, quant, lf, rp, cx, cb, lff)
            score_out[0] = orc.ssimulacra2(lin, olin)[0]
            return 0
        except Exception:  # noqa: BLE001 - reported as a status to the navigator
            return -1

    out = Outcome()
    cb_fn = PROBE_FN(probe)
    rc = lib().cc_navigate_callback(target, effort, max_probes, pred, prior, mode, cb_fn, ctypes.byref(out))
    if rc:
        raise RuntimeError(f"CPU navigator failed ({rc})")
    plan, _, _, _ = plan_at(out.rung, bool(out.chosen_fresh))
    return out, plan
