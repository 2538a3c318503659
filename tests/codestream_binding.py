"""Test infrastructure for the codestream stage: builds the checker
(tests/host/codestream_check.cpp, a restatement of the reference decoder's
reading rules that returns the quantized plan a stream carries), wraps it and
the library's jpxlg_emit_codestream, and restates the QuantizedFrameIr gather
(jpxl-encode/src/vardct/plan.rs:434-636) so oracle plans can be emitted on
the CPU.  Never shipped."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "codestream_check.cpp")
SO = os.path.join(ROOT, "tests", "host", "libcodestream_check.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(SO) or os.path.getmtime(SRC) > os.path.getmtime(SO):
        subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-o", SO, SRC], check=True)
    return SO


class Info(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("width", "height", "bits", "xyb", "encoding", "frame_type",
                                                 "x_qm_scale", "b_qm_scale", "gaborish", "epf_iters",
                                                 "num_passes", "is_last")] + [
        ("flags", ctypes.c_uint64), ("global_scale", ctypes.c_uint32), ("quant_lf", ctypes.c_uint32),
        ("x_factor_lf", ctypes.c_int32), ("b_factor_lf", ctypes.c_int32), ("lf_corr_default", ctypes.c_uint32),
        ("num_sections", ctypes.c_uint32), ("num_groups", ctypes.c_uint32), ("num_lf_groups", ctypes.c_uint32),
        ("header_bytes", ctypes.c_uint64), ("toc_sum", ctypes.c_uint64), ("stage", ctypes.c_uint32),
        ("colour_space", ctypes.c_uint32), ("colour_factor", ctypes.c_uint32), ("base_corr_x", ctypes.c_float),
        ("base_corr_b", ctypes.c_float), ("global_tree", ctypes.c_uint32), ("used_orders", ctypes.c_uint32),
        ("extra_precision", ctypes.c_uint32), ("dct_select_histogram", ctypes.c_uint32 * 27),
        ("m_lf", ctypes.c_float * 3), ("intensity_target", ctypes.c_float)]


class Varblock(ctypes.Structure):
    _fields_ = [("dct_select", ctypes.c_uint32), ("bx", ctypes.c_uint32), ("by", ctypes.c_uint32),
                ("hf_mul", ctypes.c_uint32), ("start", ctypes.c_uint64)]


class Full(ctypes.Structure):
    _fields_ = [("coeffs", ctypes.c_void_p), ("varblocks", ctypes.POINTER(Varblock)), ("capacity", ctypes.c_uint32),
                ("count", ctypes.c_uint32), ("extra_precision", ctypes.c_void_p)]


_CHECK = None


def checker():
    global _CHECK
    if _CHECK is None:
        lib = ctypes.CDLL(build())
        lib.cs_decode.restype = ctypes.c_int
        lib.cs_decode.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(Info)] + [ctypes.c_void_p] * 7 + [
            ctypes.c_char_p, ctypes.c_int]
        u64p = ctypes.POINTER(ctypes.c_uint64)
        lib.cs_prefix_decode.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                                         ctypes.c_void_p, ctypes.c_int, u64p]
        lib.cs_cluster_map.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p, u64p]
        lib.cs_hybrid.argtypes = [ctypes.c_uint32] * 4 + [ctypes.c_char_p, ctypes.c_size_t,
                                                           ctypes.POINTER(ctypes.c_uint32), u64p]
        lib.cs_natural_order.argtypes = [ctypes.c_uint32, ctypes.c_void_p]
        lib.cs_decode_full.restype = ctypes.c_int
        lib.cs_decode_full.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(Info)] + [
            ctypes.c_void_p] * 5 + [ctypes.POINTER(Full), ctypes.c_char_p, ctypes.c_int]
        _CHECK = lib
    return _CHECK


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Decoded:
    def __init__(self, info, quant, lf, bmap, hf_mul, cfl_x, cfl_b, sharp):
        self.info, self.quant, self.lf, self.block_map = info, quant, lf, bmap
        self.hf_mul, self.cfl_x, self.cfl_b, self.sharpness = hf_mul, cfl_x, cfl_b, sharp


def decode(stream: bytes) -> Decoded:
    """Reads a codestream; raises ValueError with the clause on refusal."""
    lib = checker()
    info = Info()
    # header pass for the geometry
    err = ctypes.create_string_buffer(256)
    lib.cs_decode(stream, len(stream), ctypes.byref(info), None, None, None, None, None, None, None, err, 256)
    w, h = info.width, info.height
    if not w:
        raise ValueError(err.value.decode())
    bw, bh = (w + 7) // 8, (h + 7) // 8
    tw, th = (bw + 7) // 8, (bh + 7) // 8
    quant = np.zeros((3, 8 * bh, 8 * bw), np.int32)
    lf = np.zeros((3, bh, bw), np.int32)
    bmap = np.zeros((bh, bw), np.uint8)
    mul = np.zeros((bh, bw), np.uint32)
    cx = np.zeros((th, tw), np.int8)
    cb = np.zeros((th, tw), np.int8)
    sh = np.zeros((bh, bw), np.uint8)
    rc = lib.cs_decode(stream, len(stream), ctypes.byref(info), _p(quant), _p(lf), _p(bmap), _p(mul), _p(cx), _p(cb),
                       _p(sh), err, 256)
    if rc != 0:
        raise ValueError(f"stage {info.stage}: {err.value.decode()}")
    return Decoded(info, quant, lf, bmap, mul, cx, cb, sh)


class FullFrame:
    """A frame's varblocks with their full quantized coefficients."""

    def __init__(self, info, lf, bmap, cfl_x, cfl_b, sharp, varblocks, coeffs, extra_precision):
        self.info, self.lf, self.block_map = info, lf, bmap
        self.cfl_x, self.cfl_b, self.sharpness = cfl_x, cfl_b, sharp
        self.varblocks, self.coeffs, self.extra_precision = varblocks, coeffs, extra_precision


def decode_full(stream: bytes) -> FullFrame:
    """Reads a codestream keeping every varblock's coefficient arrays (any
    DctSelect type): varblocks is a list of (dct_select, bx, by, hf_mul,
    start); coeffs[start + c * n : start + (c + 1) * n] is channel c's
    landscape coefficient array (X, Y, B) of n = 64 x blocks cells."""
    lib = checker()
    info = Info()
    err = ctypes.create_string_buffer(256)
    lib.cs_decode(stream, len(stream), ctypes.byref(info), None, None, None, None, None, None, None, err, 256)
    w, h = info.width, info.height
    if not w:
        raise ValueError(err.value.decode())
    bw, bh = (w + 7) // 8, (h + 7) // 8
    tw, th = (bw + 7) // 8, (bh + 7) // 8
    lf = np.zeros((3, bh, bw), np.int32)
    bmap = np.zeros((bh, bw), np.uint8)
    cx = np.zeros((th, tw), np.int8)
    cb = np.zeros((th, tw), np.int8)
    sh = np.zeros((bh, bw), np.uint8)
    coeffs = np.zeros(3 * 64 * bw * bh, np.int32)
    table = (Varblock * (bw * bh))()
    ep = np.zeros(max(1, ((w + 2047) // 2048) * ((h + 2047) // 2048)), np.uint8)
    full = Full(coeffs.ctypes.data, table, bw * bh, 0, ep.ctypes.data)
    rc = lib.cs_decode_full(stream, len(stream), ctypes.byref(info), _p(lf), _p(bmap), _p(cx), _p(cb), _p(sh),
                            ctypes.byref(full), err, 256)
    if rc != 0:
        raise ValueError(f"stage {info.stage}: {err.value.decode()}")
    vbs = [(t.dct_select, t.bx, t.by, t.hf_mul, t.start) for t in table[:full.count]]
    return FullFrame(info, lf, bmap, cx, cb, sh, vbs, coeffs, ep)


def prefix_decode(lengths, data: bytes, count: int):
    """(symbols, bits read) of the checker's canonical prefix decode."""
    out = np.zeros(count, np.uint32)
    nb = ctypes.c_uint64()
    rc = checker().cs_prefix_decode(bytes(lengths), len(lengths), data, len(data), _p(out), count, ctypes.byref(nb))
    if rc:
        raise ValueError("prefix code refused")
    return out.tolist(), nb.value


def cluster_map(data: bytes, num_dist: int):
    out = np.zeros(num_dist, np.uint8)
    nb = ctypes.c_uint64()
    if checker().cs_cluster_map(data, len(data), num_dist, _p(out), ctypes.byref(nb)):
        raise ValueError("cluster map refused")
    return out.tolist(), nb.value


def hybrid(se, msb, lsb, token, data: bytes = b"\0" * 8):
    v, nb = ctypes.c_uint32(), ctypes.c_uint64()
    if checker().cs_hybrid(se, msb, lsb, token, data, len(data), ctypes.byref(v), ctypes.byref(nb)):
        raise ValueError("hybrid integer refused")
    return v.value, nb.value


def natural_order(side: int) -> np.ndarray:
    out = np.zeros(side * side, np.uint32)
    checker().cs_natural_order(side, _p(out))
    return out


def ir_from_footprint(quant, lf, bmap, hf_mul=None):
    """QuantizedFrameIr restated from its definition (plan.rs:434-636) over
    the footprint layout: LF groups of 256 x 256 blocks in raster order; in a
    group, varblocks by the block-raster order of their top-left blocks, each
    with its X, Y, B arrays in the coefficient array's raster order (cell k of
    a side-s square at row 8 by + k / s, column 8 bx + k % s); each group's LF
    planes in raster order."""
    bh, bw = bmap.shape
    arena, lfg, vbs = [], [], []
    start = 0
    for gy in range(0, bh, 256):
        for gx in range(0, bw, 256):
            gi = (gy // 256) * ((bw + 255) // 256) + gx // 256
            for c in range(3):
                lfg.append(lf[c, gy:gy + 256, gx:gx + 256].reshape(-1))
            for by in range(gy, min(bh, gy + 256)):
                for bx in range(gx, min(bw, gx + 256)):
                    s = int(bmap[by, bx])
                    if not s:
                        continue
                    for c in range(3):
                        arena.append(quant[c, 8 * by:8 * by + s, 8 * bx:8 * bx + s].reshape(-1))
                    mul = int(hf_mul[by, bx]) if hf_mul is not None else 1
                    vbs.append(dict(lf_group=gi, bx=bx, by=by, side=s, hf_mul=mul, start=start))
                    start += 3 * s * s
    return np.concatenate(arena).astype(np.int32), np.concatenate(lfg).astype(np.int32), vbs


def varblock_table(vbs):
    from jpxl_amd import _lib

    table = (_lib.Varblock * max(1, len(vbs)))()
    for i, v in enumerate(vbs):
        t = table[i]
        t.lf_group, t.bx, t.by, t.side = v["lf_group"], v["bx"], v["by"], v["side"]
        t.hf_mul, t.start = v.get("hf_mul", 1), v["start"]
    return table


def emit(frame, arena, lfg, vbs, cfl_x=None, cfl_b=None, threads: int = 4):
    """jpxlg_emit_codestream through the C ABI: (bytes, Sizes)."""
    from jpxl_amd import _lib

    lib = _lib.load()
    table = varblock_table(vbs)
    n = ctypes.c_size_t()
    sizes = _lib.EmitSizes()
    cx = np.ascontiguousarray(cfl_x, np.int8) if cfl_x is not None else None
    cb = np.ascontiguousarray(cfl_b, np.int8) if cfl_b is not None else None
    rc = lib.jpxlg_emit_codestream(ctypes.byref(frame), _p(arena), _p(lfg), table, len(vbs), _p(cx), _p(cb), threads,
                                   None, 0, ctypes.byref(n), ctypes.byref(sizes))
    if rc != 0:
        raise RuntimeError(f"jpxlg_emit_codestream: {rc}")
    out = np.empty(n.value, np.uint8)
    rc = lib.jpxlg_emit_codestream(ctypes.byref(frame), _p(arena), _p(lfg), table, len(vbs), _p(cx), _p(cb), threads,
                                   _p(out), out.size, ctypes.byref(n), ctypes.byref(sizes))
    if rc != 0:
        raise RuntimeError(f"jpxlg_emit_codestream: {rc}")
    return out.tobytes(), sizes


def llf_zeroed(quant, bmap):
    """The footprint planes with each varblock's LLF cells zeroed (they are
    carried by the LF planes, not coded as HF coefficients)."""
    q = quant.copy()
    bh, bw = bmap.shape
    for by in range(bh):
        for bx in range(bw):
            s = int(bmap[by, bx])
            if s:
                c = s // 8
                q[:, 8 * by:8 * by + c, 8 * bx:8 * bx + c] = 0
    return q


def runs_from_ir(arena, vbs):
    """The token-ready runs of an IR (restated from their definition): per
    varblock and channel, the HF coefficients in natural order up to the last
    non-zero one.  -> (run arena int32, list of (start, count, nonzeros))."""
    orders = {s: natural_order(s) for s in (8, 16, 32)}
    out, runs, at = [], [], 0
    for v in vbs:
        s = v["side"]
        nb = (s // 8) ** 2
        for c in range(3):
            a = arena[v["start"] + c * s * s: v["start"] + (c + 1) * s * s]
            nat = a[orders[s]][nb:]
            nzp = np.nonzero(nat)[0]
            n = int(nzp[-1]) + 1 if nzp.size else 0
            out.append(nat[:n])
            runs.append((at, n, int(nzp.size)))
            at += n
    arena_out = np.concatenate(out).astype(np.int32) if out else np.zeros(0, np.int32)
    return arena_out, runs


def run_table(runs):
    from jpxl_amd import _lib

    t = (_lib.Run * max(1, len(runs)))()
    for i, r in enumerate(runs):
        st, n, nz = (r.start, r.count, r.nonzeros) if hasattr(r, "start") else r
        t[i].start, t[i].count, t[i].nonzeros = st, n, nz
    return t


def emit_runs(frame, run_arena, runs, lfg, vbs, cfl_x=None, cfl_b=None, threads: int = 4):
    """jpxlg_emit_codestream_runs through the C ABI: (bytes, Sizes); raises
    RuntimeError(rc) on refusal."""
    from jpxl_amd import _lib

    lib = _lib.load()
    table = varblock_table(vbs)
    rt = run_table(list(runs))
    ra = np.ascontiguousarray(run_arena, np.int32)
    cx = np.ascontiguousarray(cfl_x, np.int8) if cfl_x is not None else None
    cb = np.ascontiguousarray(cfl_b, np.int8) if cfl_b is not None else None
    n = ctypes.c_size_t()
    sizes = _lib.EmitSizes()
    args = (ctypes.byref(frame), _p(ra) if ra.size else None, ra.size, rt, _p(lfg), table, len(vbs), _p(cx), _p(cb),
            threads)
    rc = lib.jpxlg_emit_codestream_runs(*args, None, 0, ctypes.byref(n), ctypes.byref(sizes))
    if rc != 0:
        raise RuntimeError(rc)
    out = np.empty(n.value, np.uint8)
    rc = lib.jpxlg_emit_codestream_runs(*args, _p(out), out.size, ctypes.byref(n), ctypes.byref(sizes))
    if rc != 0:
        raise RuntimeError(rc)
    return out.tobytes(), sizes
