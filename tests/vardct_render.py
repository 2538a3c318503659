"""Test infrastructure: the samples of a VarDCT codestream, reconstructed from
the reader's full coefficients (codestream_binding.decode_full) the way the
reference decoder renders them:

  I.5.2  LF: m * q / 2^extra_precision, chroma from luma with the LF factors,
         adaptive smoothing over the frame-wide LF image unless the frame
         skips it (jpxl-decode/src/vardct/lf.rs:286-420, decode.rs:3246-3330)
  I.5.3  HF: bias_adjust(q) * Mul * pow(0.8, qm - 2) * matrix (render.rs:248-362)
  I.6    chroma from luma with the 64 x 64 tile's factors (render.rs:364-379)
  I.8    the LLF cells from the LF rectangle (varblock.rs:817-860)
  I.9    the transform type's inverse (varblock.rs:898-915, :1229-1262)
  J      Gaborish, then EPF (render.rs:563-590)
  L.2.2  inverse XYB, then the sRGB transfer function, odd below zero
         (jpxl-core/src/color.rs:490-501, :673-701); grey frames keep plane 0
         (decode.rs:3570-3620).

The oracle's primitives carry every step they cover: the DCT8/16/32
dequantization matrices, the bias adjustment, Mul, the qm factor, the LF
multipliers, the 2-D DCTs up to 32 points, ScaleF, Gaborish, EPF and the
opsin inverse.  Only what the encoder's render never takes is restated here:
the matrices of the other transform types (I.2.4 over Table I.6's printed
parameters, extracted into tests/golden/reference_tables.json), IDCTs with a
64-point side (the f64 definition of I.2), DCT2x2 (I.9.3), the LF smoothing
and the output transfer function.  Graded against the reference's djxl
decodes, this pins those oracle primitives on reference-held vectors."""
from __future__ import annotations

import ctypes
import math

import numpy as np

import oracle_binding as orc
import reference_restated as rr

F32 = np.float32

# DctSelect -> (block rows, block columns), Table I.1 (jpxl-core/src/varblock.rs:188-251)
BLOCKS = {0: (1, 1), 2: (1, 1), 4: (2, 2), 5: (4, 4), 6: (2, 1), 7: (1, 2), 8: (4, 1), 9: (1, 4), 10: (4, 2),
          11: (2, 4), 18: (8, 8), 19: (8, 4), 20: (4, 8)}
# DctSelect -> Table I.4 parameters index
DEQUANT_INDEX = {0: 0, 2: 2, 4: 4, 5: 5, 6: 6, 7: 6, 8: 7, 9: 7, 10: 8, 11: 8, 18: 11, 19: 12, 20: 12}
# Table I.4 matrix sizes (rows, columns), landscape
MATRIX_SIZE = {0: (8, 8), 2: (8, 8), 4: (16, 16), 5: (32, 32), 6: (8, 16), 7: (8, 32), 8: (16, 32), 11: (64, 64),
               12: (32, 64)}
# J.4.3's default sharpness table and quantization multiplier (Table J.1)
EPF_SHARP_LUT = [0.0, 1 / 7, 2 / 7, 3 / 7, 4 / 7, 5 / 7, 6 / 7, 1.0]
EPF_QUANT_MUL = 0.46


def _lib():
    o = orc.lib()
    f = ctypes.c_float
    o.orc_hf_multiplier.restype = f
    o.orc_hf_multiplier.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    o.orc_qm_multiplier.restype = f
    o.orc_qm_multiplier.argtypes = [ctypes.c_int]
    o.orc_hf_recon.restype = f
    o.orc_hf_recon.argtypes = [ctypes.c_int32, f, ctypes.c_int]
    return o


def dct_quant_weights(rows: int, cols: int, params) -> np.ndarray:
    """I.2.4 mode DCT over a rows x cols landscape matrix
    (jpxl-core/src/dequant.rs:485-509)."""
    bands = rr.dq_bands(params)
    mx = math.sqrt(2.0) + 1e-6
    w = np.zeros((rows, cols), np.float64)
    for y in range(rows):
        for x in range(cols):
            dx = x / (cols - 1)
            dy = y / (rows - 1)
            w[y, x] = rr.dq_interpolate(math.sqrt(dx * dx + dy * dy), mx, bands)
    return w


def dct2_weights(params) -> np.ndarray:
    """I.2.4 mode DCT2: six rectangles and the DC (dequant.rs:1328-1356)."""
    w = np.zeros((8, 8), np.float64)

    def fill(x0, y0, x1, y1, v):
        w[y0:y1, x0:x1] = v

    fill(4, 4, 8, 8, params[5])
    fill(4, 0, 8, 4, params[4])
    fill(0, 4, 4, 8, params[4])
    fill(2, 2, 4, 4, params[3])
    fill(2, 0, 4, 2, params[2])
    fill(0, 2, 2, 4, params[2])
    fill(1, 1, 2, 2, params[1])
    fill(0, 1, 1, 2, params[0])
    fill(1, 0, 2, 1, params[0])
    w[0, 0] = 1.0
    return w


_MATRICES: dict = {}


def dequant_matrix(index: int, channel: int) -> np.ndarray:
    """The f32 reciprocal of Table I.6's default weights for a parameters
    index (DequantMatrices::matrix, dequant.rs:1233-1249); DCT8/16/32 come
    from the oracle."""
    key = (index, channel)
    if key in _MATRICES:
        return _MATRICES[key]
    if index in (0, 4, 5):
        m = orc.dequant_matrix({0: 8, 4: 16, 5: 32}[index], channel)
    else:
        t = rr.tables()["dequant_other_params"]
        rows, cols = MATRIX_SIZE[index]
        if index == 2:
            w = dct2_weights(t["dct2"][channel])
        else:
            if index in (11, 12):
                params = [t["large_bases"][index - 11][channel]] + t["seq"][channel]
            else:
                params = t[{6: "16x8", 7: "32x8", 8: "16x32"}[index]][channel]
            w = dct_quant_weights(rows, cols, params)
        m = (1.0 / w).astype(F32)
    _MATRICES[key] = m
    return m


def _idct_basis(n: int) -> np.ndarray:
    """I.2's inverse: x[k] = X[0] + sum sqrt(2) X[u] cos(pi u (k + 1/2) / n)."""
    k = np.arange(n)[:, None]
    u = np.arange(n)[None, :]
    t = np.cos(np.pi * u * (k + 0.5) / n) * math.sqrt(2.0)
    t[:, 0] = 1.0
    return t


def idct_2d(coeffs: np.ndarray, rows: int, cols: int) -> np.ndarray:
    """IDCT_2D of I.7.3: a landscape coefficient matrix to rows x cols
    samples (jpxl-core/src/dct.rs:982-997); the oracle's f32 transform when
    both sides are at most 32, else the f64 definition."""
    if max(rows, cols) <= 32:
        a = np.ascontiguousarray(coeffs, F32)
        out = np.empty((rows, cols), F32)
        orc.lib().orc_dct2d_inv(rows, cols, orc.P(a), orc.P(out))
        return out
    c = np.asarray(coeffs, np.float64)
    # rows of a strictly landscape matrix are vertical frequencies; those of a
    # square or portrait one horizontal (dct.rs:982-997 transposes first)
    m = c if cols > rows else c.T
    return (_idct_basis(rows) @ m @ _idct_basis(cols).T).astype(F32)


def aux_idct_2x2(block: np.ndarray, s: int) -> np.ndarray:
    """I.9.3 AuxIDCT2x2 on the top-left s x s cells ([y, x] indexing)."""
    out = block.copy()
    n = s // 2
    for y in range(n):
        for x in range(n):
            c00, c01 = block[y, x], block[y, n + x]
            c10, c11 = block[y + n, x], block[y + n, n + x]
            out[2 * y, 2 * x] = c00 + c01 + c10 + c11
            out[2 * y, 2 * x + 1] = c00 + c01 - c10 - c11
            out[2 * y + 1, 2 * x] = c00 - c01 + c10 - c11
            out[2 * y + 1, 2 * x + 1] = c00 - c01 - c10 + c11
    return out


def samples_from_coefficients(ds: int, coeffs: np.ndarray) -> np.ndarray:
    """I.9 for the transform types of Table I.1 this restates."""
    rows, cols = BLOCKS[ds]
    if ds == 2:
        b = coeffs.astype(F32)
        for s in (2, 4, 8):
            b = aux_idct_2x2(b, s)
        return b
    return idct_2d(coeffs, 8 * rows, 8 * cols)


def llf_from_lf(ds: int, rect: np.ndarray) -> np.ndarray:
    """I.8: the LLF cells from the varblock's LF rectangle (image
    orientation), landscape (varblock.rs:817-860)."""
    rows, cols = BLOCKS[ds]
    if ds in (0, 2):
        return rect.astype(F32).reshape(1, 1)
    cy, cx = min(rows, cols), max(rows, cols)
    inp = np.ascontiguousarray(rect.T if rows > cols else rect, F32)
    dc = np.empty((cy, cx), F32)
    o = orc.lib()
    o.orc_dct2d_fwd(cy, cx, orc.P(inp), orc.P(dc))
    sy = np.array([o.orc_scale_f(y, cy) for y in range(cy)], F32)
    sx = np.array([o.orc_scale_f(x, cx) for x in range(cx)], F32)
    return (dc * sy[:, None]) * sx[None, :]


def adaptive_smoothing(planes: np.ndarray, mult: np.ndarray) -> np.ndarray:
    """I.5.2's smoothing of the frame-wide LF image (lf.rs:351-414)."""
    _, h, w = planes.shape
    out = planes.copy()
    if w < 3 or h < 3:
        return out
    centre, edge, diag = F32(0.052262735), F32(0.2034514), F32(0.03348292)
    p = planes
    s = p[:, 1:-1, 1:-1]
    wa = (centre * s + edge * (p[:, 1:-1, :-2] + p[:, 1:-1, 2:] + p[:, :-2, 1:-1] + p[:, 2:, 1:-1])
          + diag * (p[:, :-2, :-2] + p[:, :-2, 2:] + p[:, 2:, :-2] + p[:, 2:, 2:]))
    gap = np.maximum(F32(0.5), np.abs(wa - s) / mult[:, None, None]).max(axis=0)
    factor = np.maximum(F32(3.0) - F32(4.0) * gap, F32(0.0))
    out[:, 1:-1, 1:-1] = (wa - s) * factor[None] + s
    return out


def linear_to_srgb(v: np.ndarray) -> np.ndarray:
    """The sRGB transfer function, odd below zero (color.rs:490-501)."""
    a = np.abs(v.astype(np.float64))
    e = np.where(a <= 0.0031308, 12.92 * a, 1.055 * np.power(a, 1.0 / 2.4) - 0.055)
    return (np.sign(v) * e).astype(F32)


def lf_image(f) -> np.ndarray:
    """I.5.2: the dequantized, chroma-corrected (and smoothed) LF planes."""
    info = f.info
    bh, bw = f.block_map.shape
    o = _lib()
    m = np.zeros(3, F32)
    o.orc_lf_multipliers(info.global_scale, info.quant_lf, orc.P(m))
    default = np.allclose(np.array(info.m_lf[:], np.float64), [1 / 32, 1 / 4, 1 / 2], rtol=0, atol=0)
    if not default:  # quantizer.rs:232-240 over the signalled weights
        m = np.array([65536.0 * (v / 128.0) / (info.global_scale * info.quant_lf) for v in info.m_lf], F32)
    lgx = (bw + 255) // 256
    shift = np.zeros((bh, bw), F32)
    for gy in range(0, bh, 256):
        for gx in range(0, bw, 256):
            shift[gy:gy + 256, gx:gx + 256] = F32(1 << int(f.extra_precision[(gy // 256) * lgx + gx // 256]))
    d = m[:, None, None] * f.lf.astype(F32) / shift[None]
    cf = F32(info.colour_factor)
    kx = F32(info.base_corr_x) + F32(info.x_factor_lf) / cf
    kb = F32(info.base_corr_b) + F32(info.b_factor_lf) / cf
    d[0] = d[0] + kx * d[1]
    d[2] = d[2] + kb * d[1]
    if not info.flags & 0x80:
        d = adaptive_smoothing(d, m)
    return d


def render_xyb(f, filters: bool = True) -> np.ndarray:
    """The frame's XYB planes after I.9 and, with `filters`, Annex J."""
    info = f.info
    w, h = info.width, info.height
    bh, bw = f.block_map.shape
    if info.intensity_target != 255.0:
        raise NotImplementedError("intensity targets other than 255")
    o = _lib()
    lfp = lf_image(f)
    planes = np.zeros((3, 8 * bh, 8 * bw), F32)
    sigma = np.zeros((bh, bw), F32)
    qm = [o.orc_qm_multiplier(info.x_qm_scale), F32(1.0), o.orc_qm_multiplier(info.b_qm_scale)]
    cf = F32(info.colour_factor)
    bias_cache: dict = {}

    def bias(q: np.ndarray, c: int) -> np.ndarray:
        out = np.empty(q.shape, F32)
        for v in np.unique(q):
            key = (int(v), c)
            if key not in bias_cache:
                bias_cache[key] = o.orc_hf_recon(int(v), 1.0, c)
            out[q == v] = bias_cache[key]
        return out

    for ds, bx, by, hf_mul, start in f.varblocks:
        if ds not in BLOCKS:
            raise NotImplementedError(f"DctSelect {ds}")
        rows, cols = BLOCKS[ds]
        n = 64 * rows * cols
        cy, cx = 8 * min(rows, cols), 8 * max(rows, cols)
        mul = F32(o.orc_hf_multiplier(info.global_scale, hf_mul))
        co = []
        for c in range(3):
            q = f.coeffs[start + c * n:start + (c + 1) * n].reshape(cy, cx)
            co.append((bias(q, c) * (mul * F32(qm[c]))) * dequant_matrix(DEQUANT_INDEX[ds], c))
        tile = (by // 8, bx // 8)
        kx = F32(info.base_corr_x) + F32(int(f.cfl_x[tile])) / cf
        kb = F32(info.base_corr_b) + F32(int(f.cfl_b[tile])) / cf
        co[0] = co[0] + kx * co[1]
        co[2] = co[2] + kb * co[1]
        for c in range(3):
            llf = llf_from_lf(ds, lfp[c, by:by + rows, bx:bx + cols])
            co[c][:llf.shape[0], :llf.shape[1]] = llf
            planes[c, 8 * by:8 * (by + rows), 8 * bx:8 * (bx + cols)] = samples_from_coefficients(ds, co[c])
        for j in range(rows):
            for i in range(cols):
                sh = int(f.sharpness[by + j, bx + i])
                sigma[by + j, bx + i] = (mul * F32(EPF_QUANT_MUL)) * F32(EPF_SHARP_LUT[sh])
    planes = np.ascontiguousarray(planes[:, :h, :w])
    if filters:
        if info.gaborish:
            planes = orc.gaborish(planes)
        steps = {0: [], 1: [1], 2: [1, 2], 3: [0, 1, 2]}[info.epf_iters]
        for s in steps:
            planes = orc.epf_step(s, planes, sigma)
    return planes


def render_output(f) -> np.ndarray:
    """The decoded samples as djxl writes them to .npy: (H, W, channels) f32
    in the image's colour encoding (sRGB transfer), unclipped."""
    xyb = render_xyb(f)
    _, h, w = xyb.shape
    lin = np.zeros_like(xyb)
    o = orc.lib()
    px = np.ascontiguousarray(xyb.reshape(3, -1).T)
    out = np.zeros(3, F32)
    flat = lin.reshape(3, -1)
    for i in range(px.shape[0]):
        o.orc_opsin_inverse_pixel(ctypes.c_float(px[i, 0]), ctypes.c_float(px[i, 1]), ctypes.c_float(px[i, 2]),
                                  orc.P(out))
        flat[:, i] = out
    enc = linear_to_srgb(lin)
    if f.info.colour_space == 1:
        return enc[0][:, :, None]
    return np.moveaxis(enc, 0, -1)
