"""FP64 reference DCT and the error report of the precision study.

The forward transform of the pixel path is, in exact arithmetic, the
orthonormal DCT-II of each N x N varblock scaled by 1/N, stored transposed
(cell k*N + r holds horizontal frequency k, vertical frequency r:
jpxl-core/src/dct.rs:956-975, :933-939).  This module evaluates that closed
form in float64 (the naive formula of jpxl-core/src/dct.rs:1375-1404) over a
whole frame in the footprint layout, and summarises how far a device form's
coefficients are from it, per transform size.
"""
from __future__ import annotations

import numpy as np

SIDES = (8, 16, 32)


def dct_basis(n: int) -> np.ndarray:
    """Orthonormal DCT-II basis D[k][x] in float64."""
    k = np.arange(n)[:, None]
    x = np.arange(n)[None, :]
    d = np.sqrt(2.0 / n) * np.cos(np.pi * (2 * x + 1) * k / (2 * n))
    d[0, :] = np.sqrt(1.0 / n)
    return d


def _varblocks(bmap: np.ndarray, side: int):
    by, bx = np.nonzero(bmap == side)
    return by, bx


def forward_frame_f64(xyb: np.ndarray, bmap: np.ndarray) -> np.ndarray:
    """Coefficients of every varblock of `bmap` in float64, footprint layout
    (3 x 8bh x 8bw).  Samples beyond the image repeat the last row/column
    (the gather's edge clamp, jpxl-encode-policy/src/lib.rs:3009-3057)."""
    _, h, w = xyb.shape
    bh, bw = bmap.shape
    out = np.zeros((3, 8 * bh, 8 * bw), np.float64)
    src = xyb.astype(np.float64)
    for side in SIDES:
        by, bx = _varblocks(bmap, side)
        if len(by) == 0:
            continue
        d = dct_basis(side)
        off = np.arange(side)
        ys = np.minimum(8 * by[:, None] + off[None, :], h - 1)   # (nb, side)
        xs = np.minimum(8 * bx[:, None] + off[None, :], w - 1)
        for c in range(3):
            blk = src[c][ys[:, :, None], xs[:, None, :]]          # (nb, y, x)
            # Out[k][r] = (1/N) sum_x sum_y D[k][x] D[r][y] X[y][x]
            co = np.einsum("kx,ry,byx->bkr", d, d, blk, optimize=True) / side
            rows = 8 * by[:, None, None] + off[None, :, None]
            cols = 8 * bx[:, None, None] + off[None, None, :]
            out[c][rows, cols] = co
    return out


def error_report(got: np.ndarray, want64: np.ndarray, bmap: np.ndarray) -> dict:
    """Max and RMS absolute error of `got` against the float64 coefficients,
    per transform size (over all three channels of its varblocks)."""
    rep = {}
    for side in SIDES:
        by, bx = _varblocks(bmap, side)
        if len(by) == 0:
            continue
        off = np.arange(side)
        rows = (8 * by[:, None, None] + off[None, :, None])
        cols = (8 * bx[:, None, None] + off[None, None, :])
        e = got[:, rows, cols].astype(np.float64) - want64[:, rows, cols]
        rep[str(side)] = {"max_abs": float(np.abs(e).max()), "rms": float(np.sqrt(np.mean(e * e))),
                          "varblocks": int(len(by))}
    return rep
