"""The oracle's first-guess features pinned on the CPU: the reference's own
preanalysis and source-feature tests (preanalysis.rs:577-652,
quality_features.rs:216-245) restated, an independent numpy restatement of
the 32x32 cell statistics on random frames, and closed forms of the DCT8
summary (a flat frame, single-frequency patterns whose energy lands in a
known cell) plus a numpy restatement over the oracle's pinned DCT."""
import numpy as np

import oracle_binding as orc


def test_solid_frame_is_all_flat_and_striped_one_is_not():
    """a_solid_frame_is_all_flat_and_a_striped_one_is_not (preanalysis.rs:585)."""
    f = orc.preanalysis(np.full((64, 96, 3), 90, np.uint8), 8)
    assert f["cells"] == 6 and f["share"][0] == 1.0
    assert f["edge_mean"] == 0 and f["noise_mean"] == 0 and f["hetero_frac"] == 0
    stripes = np.full((64, 96, 3), 255, np.uint8)
    stripes[:, (np.arange(96) % 4) < 2] = 0
    g = orc.preanalysis(stripes, 8)
    assert g["share"][0] == 0.0
    assert g["edge_mean"] > 0.5 and g["axis_aligned_mean"] > 0.9 and g["log2var_mean"] > 10.0


def test_block_frame_differs_from_a_solid_one():
    """a_block_frame_differs_from_a_solid_one (preanalysis.rs:615)."""
    block = np.full((128, 128, 3), 40, np.uint8)
    block[40:88, 40:88] = [250, 30, 30]
    f = orc.preanalysis(block, 8)
    s = orc.preanalysis(np.full((128, 128, 3), 40, np.uint8), 8)
    assert f["share"][0] < 1.0 and s["share"][0] == 1.0
    assert f["edge_mean"] > 0.0
    assert f["chroma_log2_q90"] > s["chroma_log2_q90"]


def test_frames_without_a_whole_cell_have_no_features():
    """frames_without_a_whole_cell_have_no_features (preanalysis.rs:635)."""
    assert orc.preanalysis(np.zeros((40, 31, 3), np.uint8), 8) is None
    assert orc.preanalysis(np.full((32, 40, 3), 1000, np.uint16), 12) is not None


def test_floor_quantiles_are_order_statistics():
    """quantiles_are_floor_indexed_order_statistics (quality_features.rs:221)."""
    v = np.array([5.0, 1.0, 4.0, 2.0, 3.0], np.float32)
    assert [float(orc.floor_quantile(v, q)) for q in (0.0, 0.5, 0.9, 1.0)] == [1.0, 3.0, 4.0, 5.0]
    assert orc.floor_quantile(np.zeros(0, np.float32), 0.5) == 0.0


def test_flat_frame_source_features():
    """a_flat_frame_is_entirely_flat_and_a_noisy_one_is_not (quality_features.rs:231)."""
    xyb, _, gray = orc.prepare_frame(np.full((48, 64, 3), 100, np.uint8), 8)
    f = orc.source_features(xyb, gray)
    assert f["flat_fraction"] == 1.0 and f["luma_variance_q90"] == 0.0 and f["grayscale"]
    noisy = ((np.arange(64 * 48 * 3) * 97 + 13) % 251).astype(np.uint8).reshape(48, 64, 3)
    xyb, _, gray = orc.prepare_frame(noisy, 8)
    f = orc.source_features(xyb, gray)
    assert f["flat_fraction"] == 0.0 and f["luma_variance_q10"] > np.float32(2e-5) and f["edge_proxy"] >= 0


def np_preanalysis(rgb8):
    """An independent vectorised restatement of preanalysis.rs:277-540."""
    h, w, _ = rgb8.shape
    cx, cy = w // 32, h // 32
    r, g, b = (rgb8[..., i].astype(np.int64) for i in range(3))
    luma = (77 * r + 150 * g + 29 * b + 128) >> 8
    cb = np.clip(((-43 * r - 85 * g + 128 * b + 128) >> 8) + 128, 0, 255)
    cr = np.clip(((128 * r - 107 * g - 21 * b + 128) >> 8) + 128, 0, 255)
    p = np.pad(luma, 1, mode="edge")
    H, W = cy * 32, cx * 32

    def at(dy, dx):
        return p[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]

    c = at(0, 0)
    tl, tc, tr, ml, mr, bl, bc, br = at(-1, -1), at(-1, 0), at(-1, 1), at(0, -1), at(0, 1), at(1, -1), at(1, 0), at(1, 1)
    gx = (tr + 2 * mr + br) - (tl + 2 * ml + bl)
    gy = (bl + 2 * bc + br) - (tl + 2 * tc + tr)
    ax, ay = np.abs(gx), np.abs(gy)
    grad = ax + ay
    edge = grad > 48
    dirbin = np.where(ax >= 2 * ay, 0, np.where(ay >= 2 * ax, 1, np.where((gx > 0) == (gy > 0), 2, 3)))
    box = (tl + tc + tr + ml + c + mr + bl + bc + br) // 9
    weak = grad < 24

    def cells(a):
        return a.reshape(cy, 32, cx, 32).sum(axis=(1, 3)).reshape(-1)

    n = 1024
    s, sq = cells(c), cells(c * c)
    mean = s // n
    var = np.maximum(sq // n - mean * mean, 0)
    cbs, cbq = cells(cb[:H, :W]), cells(cb[:H, :W] ** 2)
    crs, crq = cells(cr[:H, :W]), cells(cr[:H, :W] ** 2)
    chroma = np.minimum(np.maximum(cbq // n - (cbs // n) ** 2, 0) + np.maximum(crq // n - (crs // n) ** 2, 0), 65535)
    dirs = np.stack([cells(edge & (dirbin == k)) for k in range(4)], 1)
    etot = dirs.sum(1)
    den = np.maximum(etot, 1)
    edge_q8 = cells(edge) * 256 // n
    flat_q8 = cells(grad < 12) * 256 // n
    wk = cells(weak)
    noise = np.where(wk > 0, cells(np.where(weak, np.abs(c - box), 0)) // np.maximum(wk, 1), 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        pr = dirs / np.maximum(etot, 1)[:, None]
        hh = -np.sum(np.where(dirs > 0, pr * np.log2(np.where(dirs > 0, pr, 1.0)), 0.0), axis=1)
    ent = np.where(etot > 0, np.clip(np.round(hh / 2.0 * 256.0), 0, 256), 0).astype(np.int64)
    dom = dirs.max(1) * 256 // den
    axis = (dirs[:, 0] + dirs[:, 1]) * 256 // den
    cls = np.full(var.shape, 4)
    conds = [
        (edge_q8 >= 64) & (ent < 110) & (axis >= 160) & (noise < 7),
        (edge_q8 >= 38) & (dom >= 128),
        (noise >= 7) & (var < 64),
        (var >= 64) & (ent >= 150),
        (var < 12) & (edge_q8 < 13),
        edge_q8 < 26,
        chroma >= 96,
    ]
    classes = [6, 5, 3, 4, 0, 1, 2]
    assigned = np.zeros(var.shape, bool)
    for cond, k in zip(conds, classes):
        sel = cond & ~assigned
        cls[sel] = k
        assigned |= cond
    m = var.size
    share = np.array([np.float32(np.count_nonzero(cls == k)) / np.float32(m) for k in range(7)], np.float32)
    lv = np.log2(1.0 + var)
    lvm = lv.mean()
    has = etot > 0
    out = {"cells": m, "share": share, "log2var_mean": lvm, "log2var_std": np.sqrt(np.mean((lv - lvm) ** 2)),
           "log2var_q10": np.percentile(lv, 10), "log2var_q90": np.percentile(lv, 90),
           "edge_mean": (edge_q8 / 256.0).mean(), "edge_q90": np.percentile(edge_q8 / 256.0, 90),
           "flat_mean": (flat_q8 / 256.0).mean(), "noise_mean": noise.mean(), "noise_q90": np.percentile(noise, 90),
           "chroma_log2_mean": np.log2(1.0 + chroma).mean(),
           "chroma_log2_q90": np.percentile(np.log2(1.0 + chroma), 90),
           "orient_entropy_mean": (ent[has] / 256.0).mean() if has.any() else 0.0,
           "dir_dominance_mean": (dom[has] / 256.0).mean() if has.any() else 0.0,
           "axis_aligned_mean": (axis[has] / 256.0).mean() if has.any() else 0.0,
           "hetero_frac": np.count_nonzero(np.abs(lv - lvm) > 2.0) / m}
    return out


def _frames():
    from jpxl_amd.synth import synthetic_rgb

    rng = np.random.default_rng(3)
    yield synthetic_rgb(200, 136, 7, 8)
    yield synthetic_rgb(97, 161, 8, 8)
    yield rng.integers(0, 256, (64, 96, 3), dtype=np.uint8)
    blk = np.full((96, 128, 3), 30, np.uint8)
    blk[20:70, 10:60] = [200, 180, 20]
    blk[::7, :, :] = 255
    yield blk


def test_oracle_preanalysis_matches_a_numpy_restatement():
    for rgb in _frames():
        f = orc.preanalysis(rgb, 8)
        want = np_preanalysis(rgb)
        assert f["cells"] == want["cells"]
        assert np.array_equal(f["share"], want["share"])
        for k, v in want.items():
            if k in ("cells", "share"):
                continue
            assert abs(float(f[k]) - float(np.float32(v))) <= 2e-7 * max(1.0, abs(float(v))), (k, f[k], v)


def test_oracle_preanalysis_of_deep_samples_uses_their_top_byte():
    from jpxl_amd.synth import synthetic_rgb

    rgb16 = synthetic_rgb(160, 96, 4, 16)
    f = orc.preanalysis(rgb16, 16)
    g = orc.preanalysis((rgb16 >> 8).astype(np.uint8), 8)
    assert all(np.array_equal(f[k], g[k]) for k in g)
    rgb12 = (rgb16 >> 4).astype(np.uint16)
    h = orc.preanalysis(rgb12, 12)
    assert all(np.array_equal(h[k], g[k]) for k in g)


# ---- the DCT8 summary ----------------------------------------------------------


def test_summary_of_a_flat_frame():
    xyb = np.full((3, 40, 48), 0.25, np.float32)
    t, hist = orc.transform_summary(xyb)
    assert t["blocks"] == 30 and hist[0] == 30 and hist.sum() == 30
    assert t["ln_ac_y_mean"] == np.log(1e-30)
    assert t["ln_ac_y_q50"] == t["ln_ac_y_q99"] == -46.0 + 0.5 * 0.125
    assert t["near_zero_frac_1e3"] == 1.0 and t["near_zero_frac_1e2"] == 1.0
    assert t["high_low_ratio"] == 0.0 and t["directional_asymmetry"] == 0.0 and t["chroma_ac_ratio"] == 0.0
    assert t["dc_variance_y"] < 1e-20


def _cosine(u, v, side=8, amp=0.1):
    x = np.arange(side)
    cx = np.cos(np.pi * (2 * x + 1) * u / 16.0)
    cy = np.cos(np.pi * (2 * x + 1) * v / 16.0)
    return amp * np.outer(cy, cx)


def test_summary_places_single_frequencies():
    """A block pattern cos(u x) cos(v y): its Y AC energy lies in cell
    (row v, column u), so the band split and the row/column asymmetry are
    fixed by (u, v)."""
    for (u, v, low, high, asym) in ((2, 0, True, False, 1.0), (0, 2, True, False, 1.0), (5, 0, False, True, 1.0),
                                    (1, 4, False, True, 1.0), (4, 4, False, True, None), (1, 1, True, False, None)):
        tile = np.tile(_cosine(u, v), (3, 4)).astype(np.float32)
        xyb = np.stack([np.zeros_like(tile), 0.5 + tile, np.zeros_like(tile)])
        t, _ = orc.transform_summary(xyb)
        if low:
            assert t["high_low_ratio"] < 1e-6, (u, v, t)
        if high:
            assert t["high_low_ratio"] > 1e6, (u, v, t)
        if asym is not None:  # diagonal cells count toward neither side
            assert abs(t["directional_asymmetry"] - asym) < 1e-6, (u, v, t)
        # a pattern along x only puts its energy in columns (col > row)
        assert t["chroma_ac_ratio"] < 1e-9
        assert t["near_zero_frac_1e3"] < 1.0


def test_summary_matches_a_numpy_restatement():
    """The summary's quantities recomputed in numpy from the oracle's DCT8
    coefficients (orc.forward_frame, pinned against the closed-form DCT)."""
    from jpxl_amd.synth import synthetic_rgb

    for w, h, seed in ((203, 141, 1), (64, 64, 2)):
        xyb, _, _ = orc.prepare_frame(synthetic_rgb(w, h, seed, 8), 8)
        t, hist = orc.transform_summary(xyb)
        bw, bh = (w + 7) // 8, (h + 7) // 8
        co = orc.forward_frame(xyb, np.full((bh, bw), 8, np.uint8))
        blk = co.reshape(3, bh, 8, bw, 8).transpose(0, 1, 3, 2, 4).reshape(3, bh * bw, 64).astype(np.float64)
        cell = np.arange(64)
        row, col = cell // 8, cell % 8
        ac = cell > 0
        e = blk[1] ** 2
        ey = e[:, ac].sum(1)
        assert t["blocks"] == bh * bw
        np.testing.assert_allclose(t["ln_ac_y_mean"], np.log(ey.mean() + 1e-30), rtol=1e-12)
        lo, hi = e[:, ac & (row + col <= 2)].sum(), e[:, ac & (np.maximum(row, col) >= 4)].sum()
        np.testing.assert_allclose(t["high_low_ratio"], hi / lo, rtol=1e-10)
        rs, cs = e[:, ac & (row > col)].sum(), e[:, ac & (col > row)].sum()
        np.testing.assert_allclose(t["directional_asymmetry"], abs(rs - cs) / (rs + cs), rtol=1e-10)
        xb = (blk[0][:, ac] ** 2).sum() + (blk[2][:, ac] ** 2).sum()
        np.testing.assert_allclose(t["chroma_ac_ratio"], xb / ey.sum(), rtol=1e-10)
        mag = np.abs(blk[1][:, ac])
        assert t["near_zero_frac_1e3"] == np.count_nonzero(mag < 1e-3) / mag.size
        assert t["near_zero_frac_1e2"] == np.count_nonzero(mag < 1e-2) / mag.size
        dc = blk[1][:, 0]
        np.testing.assert_allclose(t["dc_variance_y"], np.mean(dc * dc) - np.mean(dc) ** 2, rtol=1e-8)
        bins = np.clip((np.log(ey + 1e-30) + 46.0) / 0.125, 0, 511).astype(int)
        assert np.array_equal(np.bincount(bins, minlength=512), hist)
        order = np.sort(bins)
        for q, k in ((0.5, "ln_ac_y_q50"), (0.9, "ln_ac_y_q90"), (0.99, "ln_ac_y_q99")):
            want = order[min(int(q * order.size), order.size - 1)]
            assert t[k] == -46.0 + (want + 0.5) * 0.125
