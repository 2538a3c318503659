"""Independent restatements of reference functions, written from the
reference's Rust source (cited per function, paths relative to JPXL/crates/)
in numpy / plain Python, to pin the C oracle where no golden vector reaches.

This is test infrastructure: nothing here is shipped or called by the
library.  Each function restates the reference's arithmetic in its own
operation order (f32 where the reference computes in f32, f64 where it
widens), with the reference's numeric tables read from the fixture
tests/golden/reference_tables.json (extracted mechanically from the
reference's source text), never from the oracle.
"""
from __future__ import annotations

import json
import math
import os
from fractions import Fraction

import numpy as np

F32 = np.float32
_HERE = os.path.dirname(os.path.abspath(__file__))


def tables() -> dict:
    with open(os.path.join(_HERE, "golden", "reference_tables.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------------------
# f32 arithmetic helpers
# ---------------------------------------------------------------------------

def fma32(a, b, c):
    """Correctly rounded f32 fused multiply-add, elementwise.

    a*b is exact in f64; p + c is rounded once to f64 and then to f32.  That
    double rounding differs from one rounding only when the f64 sum lands
    exactly on the midpoint of two f32 values while the exact sum does not;
    those elements are resolved from the exact error term of the sum."""
    a = np.asarray(a, F32)
    b = np.asarray(b, F32)
    c = np.asarray(c, F32)
    a, b, c = np.broadcast_arrays(a, b, c)
    p = a.astype(np.float64) * b.astype(np.float64)
    c64 = c.astype(np.float64)
    s = p + c64
    # two-sum error: s + e == p + c exactly
    bb = s - p
    e = (p - (s - bb)) + (c64 - bb)
    r = s.astype(F32)
    bits = s.view(np.uint64) & np.uint64((1 << 29) - 1)
    hazard = (bits == np.uint64(1 << 28)) & (e != 0.0) & np.isfinite(s)
    if np.any(hazard):
        r = np.array(r, copy=True)
        for idx in zip(*np.nonzero(hazard)):
            sv, ev = s[idx], e[idx]
            cand = F32(sv)
            up = np.nextafter(cand, F32(np.inf)) if float(cand) < sv else cand
            dn = cand if float(cand) < sv else np.nextafter(cand, F32(-np.inf))
            r[idx] = up if ev > 0 else dn
    return r


def round_half_away32(x):
    """f32 round with ties away from zero (Rust f32::round)."""
    x = np.asarray(x, F32)
    t = np.trunc(x)
    f = x - t
    return np.where(np.abs(f) >= F32(0.5), t + np.sign(x).astype(F32), t).astype(F32)


def powi32(a: float, n: int) -> np.float32:
    """f32 powi by repeated squaring (the compiler runtime's powi)."""
    a = F32(a)
    r = F32(1.0)
    recip = n < 0
    n = abs(n)
    while True:
        if n & 1:
            r = F32(r * a)
        n //= 2
        if n == 0:
            break
        a = F32(a * a)
    return F32(F32(1.0) / r) if recip else r


# ---------------------------------------------------------------------------
# I.2.4 dequantization weights (jpxl-core/src/dequant.rs:400-510, :321-338)
# ---------------------------------------------------------------------------

def dq_mult(v: float) -> float:
    return 1.0 + v if v > 0.0 else 1.0 / (1.0 - v)


def dq_interpolate(pos: float, mx: float, bands) -> float:
    last = len(bands) - 1
    if last == 0:
        return bands[0]
    scaled = pos * float(last) / mx
    idx = math.floor(scaled)
    frac = scaled - idx
    a, b = bands[idx], bands[idx + 1]
    return a * math.pow(b / a, frac)


def dq_bands(params) -> list[float]:
    out = [params[0]]
    for p in params[1:]:
        out.append(out[-1] * dq_mult(p))
    return out


def dequant_matrix(side: int, channel: int, tab: dict | None = None) -> np.ndarray:
    """DequantMatrices::matrix for the square DCT of `side`: the f32
    reciprocal of GetDctQuantWeights, row-major [y][x]."""
    tab = tab or tables()
    params = tab["dequant_dct_params"][str(side)][channel]
    bands = dq_bands(params)
    mx = math.sqrt(2.0) + 1e-6
    span = float(side - 1)
    out = np.zeros((side, side), F32)
    for y in range(side):
        for x in range(side):
            dx = float(x) / span
            dy = float(y) / span
            dist = math.sqrt(dx * dx + dy * dy)
            out[y, x] = F32(1.0 / dq_interpolate(dist, mx, bands))
    return out


# ---------------------------------------------------------------------------
# Quantizer-donor cell weights (jpxl-encode-policy/src/csf.rs:95-105, :221-316)
# ---------------------------------------------------------------------------

def radial_frequency(u: int, v: int, side: int) -> float:
    fu, fv, n = float(u), float(v), float(side)
    a, b = fu / n, fv / n
    return math.sqrt(a * a + b * b) / math.sqrt(2.0)


def donor_weights(side: int, llf: int, tab: dict | None = None) -> np.ndarray:
    m8 = dequant_matrix(8, 1, tab)
    pts = []
    for cell in range(64):
        u, v = cell // 8, cell % 8
        if u < 1 and v < 1:
            continue
        step = float(m8[u, v])  # matrix.at(x = v, y = u)
        pts.append((radial_frequency(u, v, 8), 1.0 / (step * step)))
    pts.sort(key=lambda p: p[0])  # stable, as the reference's sort_by

    def sample(f):
        if f <= pts[0][0]:
            return pts[0][1]
        if f >= pts[-1][0]:
            return pts[-1][1]
        for (f0, w0), (f1, w1) in zip(pts, pts[1:]):
            if f1 >= f:
                if f1 == f0:
                    return w1
                return w0 + (w1 - w0) * ((f - f0) / (f1 - f0))
        return pts[-1][1]

    cells = side * side
    w = np.ones(cells, F32)
    raw = np.zeros(cells)
    total, count = 0.0, 0
    for cell in range(cells):
        u, v = cell // side, cell % side
        if u < llf and v < llf:
            continue
        raw[cell] = sample(radial_frequency(u, v, side))
        total += raw[cell]
        count += 1
    mean = total / count
    for cell in range(cells):
        u, v = cell // side, cell % side
        if u < llf and v < llf:
            continue
        w[cell] = F32(raw[cell] / mean)
    return w


# ---------------------------------------------------------------------------
# HF quantizer (jpxl-encode-policy/src/quantize.rs:301-366, :1400-1499,
# :1720-1746; jpxl-core/src/reconstruct.rs:27-60; lib.rs:2109-2172)
# ---------------------------------------------------------------------------

MAX_QUANT = 1 << 20


def quant_bias(tab: dict | None = None):
    tab = tab or tables()
    out = []
    for e in tab["quant_bias_exprs"]:
        a, b = [s.strip() for s in e.split("-")]
        out.append(F32(F32(float(a)) - F32(float(b))))
    return out, F32(tab["quant_bias_numerator"])


def hf_multiplier(gs: int, mul: int) -> np.float32:
    denom = float(gs) * float(mul)
    return F32(65536.0 / denom) if denom > 0.0 else F32(0.0)


def qm_multiplier(k: int) -> np.float32:
    return powi32(0.8, k - 2)


class HfQuantizer:
    """One (side, global_scale, HfMul, qm scales) quantizer with the cover's
    Lagrange weights installed (lambda from the frame's DCT8 baseline)."""

    def __init__(self, side: int, gs: int, mul: int, xqm: int, bqm: int, lambda_scale: float,
                 baseline_mul: int | None = None, donor: bool = True, tab: dict | None = None):
        tab = tab or tables()
        self.side = side
        self.n = side // 8
        self.bias, self.numer = quant_bias(tab)
        m = hf_multiplier(gs, mul)
        scale = [F32(m * qm_multiplier(xqm)), m, F32(m * qm_multiplier(bqm))]
        self.step = np.zeros((3, side * side), F32)
        self.thr = np.zeros((3, side * side), F32)
        for c in range(3):
            mat = dequant_matrix(side, c, tab)
            for cell in range(side * side):
                x, y = cell % side, cell // side
                st = F32(scale[c] * mat[y, x])
                self.step[c, cell] = st
                self.thr[c, cell] = F32(F32(F32(0.5) * self.bias[c]) * st)
        # lambda: from the baseline DCT8 quantizer's steps
        bm = mul if baseline_mul is None else baseline_mul
        b = hf_multiplier(gs, bm)
        bscale = [F32(b * qm_multiplier(xqm)), b, F32(b * qm_multiplier(bqm))]
        m8 = [dequant_matrix(8, c, tab) for c in range(3)]
        sc = float(lambda_scale) if (math.isfinite(lambda_scale) and lambda_scale > 0) else 1.0
        self.lam = []
        for c in range(3):
            total = 0.0
            for cell in range(1, 64):
                st = 8.0 * float(F32(bscale[c] * m8[c][cell // 8, cell % 8]))
                total += st * st
            mean = total / 63.0
            self.lam.append(sc * 16.0 / mean if mean > 0 else 0.0)
        self.freq = donor_weights(side, self.n, tab) if donor else np.ones(side * side, F32)
        tsd = float(side * side)
        self.weight = np.array([[F32(self.lam[c] * tsd * float(self.freq[cell])) for cell in range(side * side)]
                                for c in range(3)], F32)

    def bias_adjust(self, q: int, c: int) -> np.float32:
        f = F32(q)
        if abs(q) <= 1:
            return F32(f * self.bias[c])
        return F32(f - F32(self.numer / f))

    def recon(self, q: int, c: int, cell: int) -> np.float32:
        return F32(self.bias_adjust(q, c) * self.step[c, cell])

    def choose(self, target, c: int, cell: int) -> int:
        target = F32(target)
        step = self.step[c, cell]
        if abs(target) <= self.thr[c, cell]:
            return 0
        v = F32(target / step)
        adj = F32(v - F32(0.5)) if np.signbit(v) else F32(v + F32(0.5))
        e = int(adj)  # truncation toward zero
        if abs(e) > MAX_QUANT:
            raise ValueError("outside the quantizer's range")
        best, best_err = 0, F32(np.inf)
        for q in (0, e - 1, e, e + 1):
            if abs(q) > MAX_QUANT:
                continue
            err = F32(abs(F32(self.bias_adjust(q, c) * step) - target))
            if err < best_err or (err == best_err and abs(q) < abs(best)):
                best, best_err = q, err
        return best


def residual_bits(q: int) -> int:
    """lib.rs residual_bits: zero is free, else the magnitude's bit length
    plus a sign bit."""
    return 0 if q == 0 else abs(int(q)).bit_length() + 1


# ---------------------------------------------------------------------------
# Trailing truncation (jpxl-encode-policy/src/quantize.rs:468-545)
# ---------------------------------------------------------------------------

def truncate_trailing(hq: HfQuantizer, c: int, quant, order, zero_token_bits: float, targets):
    quant = [int(v) for v in quant]
    num_blocks = hq.n * hq.n
    size = len(order)
    ztb = F32(zero_token_bits)

    def prev_nonzero(before):
        for pos in range(before - 1, num_blocks - 1, -1):
            if quant[order[pos]] != 0:
                return pos
        return None

    last = prev_nonzero(size)
    while last is not None:
        cell = order[last]
        q = quant[cell]
        if q == 0:
            last = prev_nonzero(last)
            continue
        weight = hq.weight[c, cell]
        previous = prev_nonzero(last)
        exposed = (last - previous - 1) if previous is not None else (last - num_blocks)
        zeros = F32(max(exposed, 0))
        own = F32(abs(q).bit_length() + 1)
        saved = F32(own + F32(zeros * ztb))
        target = F32(targets[cell])
        recon = hq.recon(q, c, cell)
        kept = F32(recon - target)
        added = fma32(target, target, -F32(kept * kept))
        cost = F32(weight * added)
        if saved <= cost:
            break
        quant[cell] = 0
        last = previous
    return quant


# ---------------------------------------------------------------------------
# CfL regression (jpxl-encode-policy/src/quantize.rs:1599-1717)
# ---------------------------------------------------------------------------

def cfl_multiplier(base: float, factor: int) -> np.float32:
    return F32(F32(base) + F32(F32(factor) / F32(84)))


def cfl_energy(sums, base, factor) -> float:
    s_yy, s_yc, s_cc = sums
    k = float(cfl_multiplier(base, factor))
    return s_cc - 2.0 * k * s_yc + k * k * s_yy


def cfl_sums(ys, cs):
    s_yy = s_yc = s_cc = 0.0
    for y, c in zip(ys, cs):
        y, c = float(F32(y)), float(F32(c))
        s_yy += y * y
        s_yc += y * c
        s_cc += c * c
    return s_yy, s_yc, s_cc


def cfl_brute_force(ys, cs, base: float, lo: int, hi: int) -> int:
    """The argmin over every integer factor of [lo, hi] of the residual
    energy, ties to the smaller magnitude (the parabola's integer minimum)."""
    sums = cfl_sums(ys, cs)
    if sums[0] <= 0.0:
        return 0
    best, best_e = None, None
    for f in sorted(range(lo, hi + 1), key=lambda v: (abs(v), v)):
        e = cfl_energy(sums, base, f)
        if best is None or e < best_e:
            best, best_e = f, e
    return best


# ---------------------------------------------------------------------------
# J.4 edge-preserving filter (jpxl-core/src/restoration.rs:56-71, :201-431,
# :583-731)
# ---------------------------------------------------------------------------

_CROSS = [(0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)]
_K13 = [(0, 0), (-1, 0), (1, 0), (0, -1), (0, 1), (1, -1), (1, 1), (-1, 1), (-1, -1), (-2, 0), (2, 0),
        (0, 2), (0, -2)]


def mirror1d(c: int, size: int) -> int:
    while True:
        if c < 0:
            c = -c - 1
        elif c >= size:
            c = 2 * size - 1 - c
        else:
            return c


def epf_step(step: int, planes: np.ndarray, sigma: np.ndarray, tab: dict | None = None) -> np.ndarray:
    """One J.4 step (0, 1, 2) over planes (3, h, w) f32 with the per-block
    sigma grid (ceil(h/8), ceil(w/8))."""
    tab = tab or tables()
    e = tab["epf"]
    planes = np.asarray(planes, F32)
    _, h, w = planes.shape
    pad = 3
    ys = np.array([mirror1d(y, h) for y in range(-pad, h + pad)])
    xs = np.array([mirror1d(x, w) for x in range(-pad, w + pad)])
    P = planes[:, ys][:, :, xs]

    def at(c, dx, dy):
        return P[c, pad + dy:pad + dy + h, pad + dx:pad + dx + w]

    taps = _K13 if step == 0 else _CROSS
    cross = [(0, 0)] if step == 2 else _CROSS
    sscale = {0: e["pass0_sigma_scale"], 1: 1.0, 2: e["pass2_sigma_scale"]}[step]
    step_mul = F32(F32(e["step_multiplier_base"]) * F32(sscale))
    num, den = [float(t) for t in e["border_sad_mul_expr"].split("/")]
    border_mul = F32(F32(num) / F32(den))
    chs = [F32(v) for v in e["channel_scale"]]
    yy, xx = np.mgrid[0:h, 0:w]
    at_border = (yy % 8 == 0) | (yy % 8 == 7) | (xx % 8 == 0) | (xx % 8 == 7)
    pm = np.where(at_border, border_mul, F32(1.0)).astype(F32)
    s = np.asarray(sigma, F32)[yy // 8, xx // 8]
    safe = np.where(s < F32(e["sigma_skip"]), F32(1.0), s).astype(F32)
    inv_sigma = (F32(F32(step_mul * F32(4.0)) * F32(F32(1.0) - np.sqrt(F32(0.5)))) / safe).astype(F32)
    sum_w = np.zeros((h, w), F32)
    acc = np.zeros((3, h, w), F32)
    for kx, ky in taps:
        d = np.zeros((h, w), F32)
        for c in range(3):
            for ix, iy in cross:
                a = at(c, ix, iy)
                b = at(c, kx + ix, ky + iy)
                d = fma32(np.abs((a - b).astype(F32)), chs[c], d)
        sd = (pm * d).astype(F32)
        v = fma32(sd, -inv_sigma, F32(1.0))
        wgt = np.where(sd == F32(0.0), F32(1.0), np.where(v > F32(0.0), v, F32(0.0))).astype(F32)
        sum_w = (sum_w + wgt).astype(F32)
        for c in range(3):
            acc[c] = fma32(at(c, kx, ky), wgt, acc[c])
    out = (acc / sum_w).astype(F32)
    skip = s < F32(e["sigma_skip"])
    out[:, skip] = planes[:, skip]
    return out


# ---------------------------------------------------------------------------
# SSIMULACRA2 pooling (jpxl-perceptual/src/pool.rs:60-68, :243-270;
# ssimulacra2.rs:509-540)
# ---------------------------------------------------------------------------

def map_sums(i1, m1, v11, i2, m2, v22, v12, c2: float) -> list[float]:
    c2 = F32(c2)
    sums = [0.0] * 6
    for k in range(len(i1)):
        a1, b1, s11 = F32(i1[k]), F32(m1[k]), F32(v11[k])
        a2, b2, s22, s12 = F32(i2[k]), F32(m2[k]), F32(v22[k]), F32(v12[k])
        mu11, mu22, mu12 = F32(b1 * b1), F32(b2 * b2), F32(b1 * b2)
        md = F32(b1 - b2)
        num_m = F32(F32(1.0) - F32(md * md))
        num_s = F32(F32(F32(2.0) * F32(s12 - mu12)) + c2)
        den_s = F32(F32(F32(s11 - mu11) + F32(s22 - mu22)) + c2)
        d = max(1.0 - float(F32(F32(num_m * num_s) / den_s)), 0.0)
        sums[0] += d
        sums[1] += d * d * d * d
        d1 = (1.0 + float(abs(F32(a2 - b2)))) / (1.0 + float(abs(F32(a1 - b1)))) - 1.0
        art = max(d1, 0.0)
        sums[2] += art
        sums[3] += art * art * art * art
        lost = max(-d1, 0.0)
        sums[4] += lost
        sums[5] += lost * lost * lost * lost
    return sums


def from_sums(sums, pixels: int) -> list[float]:
    inv = 1.0 / max(pixels, 1)
    ones = [inv * sums[2 * m] for m in range(3)]
    fours = [math.sqrt(math.sqrt(inv * sums[2 * m + 1])) for m in range(3)]
    return ones + fours


def weighted_error(terms, nscales: int, tab: dict | None = None) -> float:
    """terms in (component, scale, norm, map) order for nscales scales."""
    tab = tab or tables()
    wts = iter(tab["ssimulacra2"]["weights"])
    err = 0.0
    for c in range(3):
        for s in range(nscales):
            for k in range(6):
                err += next(wts) * abs(terms[c * nscales * 6 + s * 6 + k])
    return err


# ---------------------------------------------------------------------------
# Output depth quantizer (jpxl-plan-render/src/lib.rs:153-216, :333-353;
# jpxl-core/src/color.rs:490-528)
# ---------------------------------------------------------------------------

_LIBM = None


def powf(a, b):
    """The C library's single-precision pow, elementwise (what Rust's
    f32::powf calls; numpy's own vectorised pow may differ in the last ulp)."""
    global _LIBM
    if _LIBM is None:
        import ctypes
        import ctypes.util

        _LIBM = ctypes.CDLL(ctypes.util.find_library("m"))
        _LIBM.powf.restype = ctypes.c_float
        _LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    a = np.asarray(a, F32)
    out = np.empty(a.shape, F32)
    fa, fo = a.reshape(-1), out.reshape(-1)
    bb = float(F32(b))
    for i in range(fa.size):
        fo[i] = _LIBM.powf(float(fa[i]), bb)
    return out


def linear_to_srgb32(v):
    v = np.asarray(v, F32)
    a = np.abs(v)
    lin = (F32(12.92) * a).astype(F32)
    curve = (F32(1.055) * powf(a, F32(F32(1.0) / F32(2.4)))).astype(F32) - F32(0.055)
    enc = np.where(a <= F32(0.0031308), lin, curve.astype(F32)).astype(F32)
    return np.where(v < 0, -enc, enc).astype(F32)


def srgb_to_linear32(v):
    v = np.asarray(v, F32)
    a = np.abs(v)
    lin = (a / F32(12.92)).astype(F32)
    curve = powf(((a + F32(0.055)) / F32(1.055)).astype(F32), F32(2.4)).astype(F32)
    dec = np.where(a <= F32(F32(12.92) * F32(0.0031308)), lin, curve).astype(F32)
    return np.where(v < 0, -dec, dec).astype(F32)


def composed_level(x, bits: int):
    mx = F32((1 << bits) - 1)
    scaled = round_half_away32((linear_to_srgb32(x) * mx).astype(F32))
    ok = np.isfinite(scaled)
    q = np.clip(np.where(ok, scaled, 0), 0, mx).astype(np.int64)
    return q


# ---------------------------------------------------------------------------
# Hierarchical cover (jpxl-encode-policy/src/lib.rs:3009-3057, :4685-4800,
# :4909-5062, :5115-5223; request.rs:585-636)
# ---------------------------------------------------------------------------

def cover_rate_model(calibrated: bool):
    if calibrated:
        return {8: 1.56, 16: 1.63, 32: 1.88}, {8: 5.5, 16: 8.4 + 4.0, 32: 6.5 + 4.0}
    return {8: 1.0, 16: 1.0, 32: 1.0}, {8: 2.0, 16: 34.0, 32: 34.0}


def gather_square(plane: np.ndarray, x0: int, y0: int, side: int) -> np.ndarray:
    h, w = plane.shape
    ys = np.minimum(np.arange(y0, y0 + side), h - 1)
    xs = np.minimum(np.arange(x0, x0 + side), w - 1)
    return plane[ys][:, xs]


def cover_brute_force(xyb: np.ndarray, gs: int, mul: int, xqm: int, bqm: int, lambda_scale: float,
                      dct_fwd, calibrated: bool = True, donor: bool = True) -> np.ndarray:
    """The block map the quadtree objective J picks, with every candidate's
    J evaluated in full (no pruning); a merge must be strictly cheaper than
    its split.  dct_fwd(block) is the forward DCT_2D of one square (the
    separately pinned transform); AQ off."""
    _, h, w = xyb.shape
    bw, bh = (w + 7) // 8, (h + 7) // 8
    scales, fixed = cover_rate_model(calibrated)
    hq = {s: HfQuantizer(s, gs, mul, xqm, bqm, lambda_scale, donor=donor) for s in (8, 16, 32)}

    def cost(side, bx, by):
        q = hq[side]
        n = side // 8
        co = [dct_fwd(gather_square(xyb[c], 8 * bx, 8 * by, side)).reshape(-1) for c in range(3)]
        bits, sse = 0, 0.0
        dy = np.zeros(side * side, F32)
        tsd = float(side * side)
        for c in (1, 0, 2):
            lw = q.lam[c] * tsd
            for row in range(side):
                for col in range(n if row < n else 0, side):
                    cell = row * side + col
                    if c == 1:
                        t = F32(co[1][cell])
                    else:
                        k = F32(0.0) if c == 0 else F32(1.0)
                        t = F32(co[c][cell] - F32(k * dy[cell]))
                    v = q.choose(t, c, cell)
                    r = q.recon(v, c, cell)
                    if c == 1:
                        dy[cell] = r
                    bits += residual_bits(v)
                    d = float(F32(r - t))
                    sse += lw * float(q.freq[cell]) * (d * d)
        return bits * scales[side] + sse

    bmap = np.zeros((bh, bw), np.uint8)

    def region(bx, by, size):
        """(cost, [(bx, by, side)]) of the best cover of the region."""
        if bx >= bw or by >= bh:
            return 0.0, []
        if size == 1:
            return cost(8, bx, by) + fixed[8], [(bx, by, 8)]
        half = size // 2
        split, blocks = 0.0, []
        for qx, qy in ((bx, by), (bx + half, by), (bx, by + half), (bx + half, by + half)):
            c, b = region(qx, qy, half)
            split += c
            blocks += b
        if bx + size <= bw and by + size <= bh:
            side = 8 * size
            single = cost(side, bx, by) + fixed[side]
            if single < split:
                return single, [(bx, by, side)]
        return split, blocks

    for sby in range(0, bh, 4):
        for sbx in range(0, bw, 4):
            for bx, by, side in region(sbx, sby, 4)[1]:
                bmap[by, bx] = side
    return bmap
