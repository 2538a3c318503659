"""Pins the oracle's functions that no golden vector of the reference reaches:
each is compared with an independent restatement written from the
reference's source (tests/reference_restated.py), with the reference's own
known-answer tests, or with what the mathematics fixes (brute force over
every candidate, closed forms).  The restatements read the reference's
numeric tables from tests/golden/reference_tables.json, which
tests/golden/extract_reference_tables.py extracted from the reference's
source text, so a transcription slip in the oracle's tables fails too.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_binding as orc
import reference_restated as rr

F32 = np.float32
c_d = ctypes.c_double


def _o():
    o = orc.lib()
    o.orc_dequant_mult.restype = c_d
    o.orc_dequant_mult.argtypes = [c_d]
    o.orc_dequant_interpolate.restype = c_d
    o.orc_dequant_interpolate.argtypes = [c_d, c_d, ctypes.c_void_p, ctypes.c_int]
    o.orc_dequant_bands.restype = ctypes.c_int
    o.orc_weighted_error.restype = c_d
    o.orc_weighted_error.argtypes = [ctypes.c_void_p, ctypes.c_int]
    o.orc_cfl_best_factor.restype = ctypes.c_int
    o.orc_cfl_best_factor.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                      ctypes.c_int, ctypes.c_int]
    o.orc_pool_maps.argtypes = [ctypes.c_size_t] + [ctypes.c_void_p] * 8
    o.orc_from_sums.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    o.orc_truncate_block.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return o


# ---------------------------------------------------------------------------
# The fixed-point helpers of the restatement itself
# ---------------------------------------------------------------------------

def test_fma32_is_one_rounding():
    # 1 + 2^-24 is a tie between 1 and 1 + 2^-23 in f32; a product slightly
    # above the tie must round up (a double rounding through f64 would not).
    a, b = F32(1.0 + 2.0 ** -23), F32(1.0 + 2.0 ** -23)
    c = F32(-(2.0 ** -23))
    exact = Fraction_f32(a, b, c)
    assert rr.fma32(a, b, c) == exact
    rng = np.random.default_rng(0)
    x, y, z = (rng.standard_normal(2000).astype(F32) for _ in range(3))
    got = rr.fma32(x, y, z)
    for i in range(0, 2000, 97):
        assert got[i] == Fraction_f32(x[i], y[i], z[i])


def Fraction_f32(a, b, c):
    """Exact a*b + c rounded once to f32 (round half to even)."""
    from fractions import Fraction

    v = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    f = F32(float(v))
    lo = f if Fraction(float(f)) <= v else np.nextafter(f, F32(-np.inf))
    hi = np.nextafter(lo, F32(np.inf))
    dl, dh = v - Fraction(float(lo)), Fraction(float(hi)) - v
    if dl < dh:
        return lo
    if dh < dl:
        return hi
    return lo if (lo.view(np.uint32) & 1) == 0 else hi


# ---------------------------------------------------------------------------
# Dequantization matrices (dequant.rs:400-510, KATs :1540-1600)
# ---------------------------------------------------------------------------

def _bands(params):
    o = _o()
    p = np.asarray(params, np.float64)
    out = np.zeros(len(p))
    rc = o.orc_dequant_bands(orc.P(p), len(p), orc.P(out))
    return rc, out


def test_dequant_mult_kats():
    o = _o()
    assert o.orc_dequant_mult(1.0) == 2.0
    assert o.orc_dequant_mult(0.0) == 1.0
    assert o.orc_dequant_mult(-1.0) == 0.5
    assert abs(o.orc_dequant_mult(-0.25) - 0.8) < 1e-15
    for v in [-1e30, -1.0, -0.5, 0.0, 0.5, 1.0, 1e30]:
        assert o.orc_dequant_mult(v) > 0.0


def test_dequant_interpolate_kats():
    o = _o()

    def it(pos, mx, bands):
        b = np.asarray(bands, np.float64)
        return o.orc_dequant_interpolate(pos, mx, orc.P(b), len(b))

    assert abs(it(0.5, 1.0, [2.0, 8.0]) - 4.0) < 1e-12  # pow(B/A, frac), not B/4B
    assert abs(it(0.0, 1.0, [2.0, 8.0]) - 2.0) < 1e-12
    assert abs(it(0.5, 1.0, [1.0, 3.0, 9.0]) - 3.0) < 1e-12
    assert it(123.0, 1.0, [7.0]) == 7.0
    # positions past the range are rejected (the +1e-6 of the callers)
    assert math.isnan(it(1.0, 1.0, [2.0, 8.0]))
    assert math.isnan(it(-1.0, 1.0, [2.0, 8.0]))


def test_dequant_bands_kats():
    rc, b = _bands([22.0, -0.25, -0.25, -0.25])
    assert rc == 0
    assert abs(b[1] - 17.6) < 1e-12 and abs(b[2] - 14.08) < 1e-12 and abs(b[3] - 11.264) < 1e-12
    assert _bands([0.0, 1.0])[0] != 0 and _bands([-3.0])[0] != 0


@pytest.mark.parametrize("side", [8, 16, 32])
@pytest.mark.parametrize("channel", [0, 1, 2])
def test_dequant_matrix_equals_the_table_restatement(side, channel):
    """Every cell of the oracle's matrix equals I.2.4 evaluated from the
    reference's Table I.6 row (a slipped digit or a swapped row fails)."""
    want = rr.dequant_matrix(side, channel)
    got = orc.dequant_matrix(side, channel)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the DC weight is the printed base itself
    base = rr.tables()["dequant_dct_params"][str(side)][channel][0]
    assert got[0, 0] == F32(1.0 / base)


def test_a_swapped_table_row_would_fail_the_matrix_check():
    tab = rr.tables()
    rows = tab["dequant_dct_params"]["16"]
    rows[0], rows[1] = rows[1], rows[0]
    assert not np.array_equal(rr.dequant_matrix(16, 0, tab), orc.dequant_matrix(16, 0))


@pytest.mark.parametrize("side", [8, 16, 32])
def test_donor_weights_equal_the_restatement(side):
    want = rr.donor_weights(side, side // 8)
    got = orc.donor_weights(side)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# ---------------------------------------------------------------------------
# Trailing truncation (quantize.rs:468-545)
# ---------------------------------------------------------------------------

def _qparams(gs, mul, xqm, bqm, ls=4.0, ztb=1.0):
    return orc.QuantParams(gs, mul, 4, xqm, bqm, 1, ls, ztb)


@pytest.mark.parametrize("side", [8, 16, 32])
@pytest.mark.parametrize("channel", [0, 2])
def test_truncation_walk_equals_the_restatement(side, channel):
    """The backward walk on random blocks (sparse levels, mixed targets): the
    oracle's decisions equal the restated walk's, cell for cell."""
    o = _o()
    rng = np.random.default_rng(side * 10 + channel)
    order = [int(v) for v in orc.natural_order(side)]
    changed = 0
    for trial in range(12):
        gs = int(rng.choice([600, 2000, 9000]))
        ztb = float(rng.choice([0.5, 1.0, 2.0]))
        qp = _qparams(gs, 1, 3, 3, 4.0, ztb)
        hq = rr.HfQuantizer(side, gs, 1, 3, 3, 4.0)
        n = side * side
        targets = (rng.standard_normal(n) * rng.choice([0.002, 0.01, 0.05])).astype(F32)
        quant = np.zeros(n, np.int32)
        for cell in range(n):
            if rng.random() < 0.3:
                quant[cell] = hq.choose(targets[cell], channel, cell)
        want = rr.truncate_trailing(hq, channel, quant, order, ztb, targets)
        got = quant.copy()
        o.orc_truncate_block(side, channel, ctypes.byref(qp), orc.P(got), orc.P(targets))
        assert list(got) == want, (trial, gs, ztb)
        changed += int(np.count_nonzero(got != quant))
    assert changed > 0  # the walk did drop coefficients


def test_truncation_invariants_by_brute_force():
    """What the walk must leave: its last kept nonzero fails the drop test,
    every dropped one passed it when it was last, and nothing before the last
    kept nonzero moved."""
    rng = np.random.default_rng(7)
    side, c = 8, 0
    order = [int(v) for v in orc.natural_order(side)]
    hq = rr.HfQuantizer(side, 1500, 1, 3, 3, 4.0)
    for _ in range(40):
        t = (rng.standard_normal(64) * 0.02).astype(F32)
        q = [hq.choose(t[k], c, k) if rng.random() < 0.5 else 0 for k in range(64)]
        out = rr.truncate_trailing(hq, c, q, order, 1.0, t)
        pos = [p for p in range(1, 64) if out[order[p]] != 0]
        last = pos[-1] if pos else None
        for p in range(1, 64):
            cell = order[p]
            if last is not None and p <= last:
                assert out[cell] == q[cell]  # untouched before the new end
            elif q[cell] != 0:
                assert out[cell] == 0


# ---------------------------------------------------------------------------
# CfL (quantize.rs:1599-1717, tests :2472-2496)
# ---------------------------------------------------------------------------

def _best(ys, cs, base, lo=-128, hi=127):
    o = _o()
    y = np.asarray(ys, F32)
    c = np.asarray(cs, F32)
    return o.orc_cfl_best_factor(len(y), orc.P(y), orc.P(c), float(base), lo, hi)


def test_cfl_reference_kats():
    ys = [-2.0, -1.0, 0.5, 3.0]
    assert _best(ys, [0.5 * v for v in ys], 0.0) == 42
    assert rr.cfl_multiplier(0.0, 42) == F32(0.5)
    assert _best([0.0, 0.0, 0.0], [-3.0, 0.0, 8.0], 1.0) == 0


def test_cfl_best_factor_is_the_parabola_minimum():
    """Against every integer factor in range (the parabola's integer argmin,
    ties to the smaller magnitude), on random regressions and clamped ranges."""
    rng = np.random.default_rng(3)
    for trial in range(300):
        n = int(rng.integers(1, 40))
        ys = (rng.standard_normal(n) * rng.choice([0.01, 1.0, 30.0])).astype(F32)
        slope = rng.uniform(-2.0, 2.0)
        cs = (slope * ys + rng.standard_normal(n) * rng.choice([0.0, 0.1, 1.0])).astype(F32)
        base = float(rng.choice([0.0, 1.0]))
        lo, hi = (-128, 127) if trial % 3 else (int(rng.integers(-20, 0)), int(rng.integers(0, 20)))
        assert _best(ys, cs, base, lo, hi) == rr.cfl_brute_force(ys, cs, base, lo, hi), trial


def test_grayscale_frame_gives_neutral_cfl():
    """A grayscale frame (R = G = B) degenerates to the neutral factors
    exactly (quantize.rs:1646-1650)."""
    w, h = 96, 80
    rng = np.random.default_rng(5)
    g = rng.integers(0, 256, (h, w), np.uint8)
    rgb = np.repeat(g[:, :, None], 3, axis=2)
    xyb, _, gray = orc.prepare_frame(rgb, 8)
    assert gray
    bmap = orc.mixed_map(w, h, seed=2)
    coeff = orc.forward_frame(xyb, bmap)
    cx, cb, lf = orc.estimate_cfl(w, h, bmap, coeff, _qparams(2000, 1, 3, 3))
    assert not cx.any() and not cb.any() and lf == (0, 0)


# ---------------------------------------------------------------------------
# EPF (restoration.rs:583-731)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("step", [0, 1, 2])
@pytest.mark.parametrize("wh", [(24, 16), (21, 13), (40, 33)])
def test_epf_step_equals_the_per_pixel_restatement(step, wh):
    w, h = wh
    rng = np.random.default_rng(step * 100 + w)
    planes = rng.random((3, h, w)).astype(F32) * F32(0.3)
    planes[:, h // 2:, :] += F32(0.2)  # an edge across blocks
    bh, bw = (h + 7) // 8, (w + 7) // 8
    sigma = rng.uniform(0.0, 2.5, (bh, bw)).astype(F32)
    sigma[0, 0] = F32(0.29)  # below the skip threshold
    got = orc.epf_step(step, planes, sigma)
    want = rr.epf_step(step, planes, sigma)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(got[:, :8, :8], planes[:, :8, :8])  # skipped block


def test_a_wrong_epf_constant_would_fail():
    tab = rr.tables()
    tab["epf"]["channel_scale"] = [40.0, 5.0, 3.0]
    w, h = 24, 16
    rng = np.random.default_rng(1)
    planes = (rng.random((3, h, w)) * 0.003).astype(F32)  # small distances: weights inside (0, 1)
    sigma = np.full((2, 3), 2.0, F32)
    assert not np.array_equal(orc.epf_step(1, planes, sigma), rr.epf_step(1, planes, sigma, tab))
    tab = rr.tables()
    tab["epf"]["border_sad_mul_expr"] = "1.0 / 1.0"
    assert not np.array_equal(orc.epf_step(1, planes, sigma), rr.epf_step(1, planes, sigma, tab))


# ---------------------------------------------------------------------------
# SSIMULACRA2 pooling and weights (pool.rs:60-68, :243-270, tests :285, :354;
# ssimulacra2.rs:391-540)
# ---------------------------------------------------------------------------

def _pool(planes7):
    o = _o()
    p = [np.ascontiguousarray(a, F32) for a in planes7]
    sums = np.zeros(6)
    o.orc_pool_maps(len(p[0]), *[orc.P(a) for a in p], orc.P(sums))
    return sums


def _from_sums(sums, n):
    o = _o()
    s = np.asarray(sums, np.float64)
    out = np.zeros(6)
    o.orc_from_sums(orc.P(s), n, orc.P(out))
    return out


def test_identical_moments_pool_to_zero():
    img = [0.2, 0.5, 0.9, 0.4]
    mu = [0.3, 0.45, 0.8, 0.5]
    s = [0.11, 0.3, 0.7, 0.2]
    sums = _pool([img, mu, s, img, mu, s, s])
    assert not sums.any()
    assert not _from_sums(sums, 4).any()


def test_four_norm_of_a_constant_map_is_the_constant():
    t = _from_sums([0.5 * 8.0, 0.0625 * 8.0, 0, 0, 0, 0], 8)
    assert abs(t[0] - 0.5) < 1e-12 and abs(t[3] - 0.5) < 1e-12


@pytest.mark.parametrize("n", [0, 1, 7, 64, 250])
def test_map_sums_equal_the_restatement(n):
    rng = np.random.default_rng(n)
    planes = [(rng.random(n) * 2 - 1).astype(F32) for _ in range(7)]
    got = _pool(planes)
    want = rr.map_sums(*planes, rr.tables()["ssim_c2"])
    assert list(got) == want


def test_every_weight_sits_at_its_term():
    """Each of the 108 (component, scale, norm, map) positions carries the
    reference's weight, and fewer scales consume the weights consecutively."""
    o = _o()
    wts = rr.tables()["ssimulacra2"]["weights"]
    for ns in (6, 4, 1):
        for k in range(18 * ns):
            t = np.zeros(108)
            t[k] = -2.0  # |term| enters the sum
            assert o.orc_weighted_error(orc.P(t), ns) == 2.0 * wts[k]
    rng = np.random.default_rng(9)
    t = rng.random(108)
    assert o.orc_weighted_error(orc.P(t), 6) == rr.weighted_error(t, 6)


def test_remap_uses_the_reference_constants():
    lits = rr.tables()["ssimulacra2"]["remap_literals"]
    a, c3, c1, c2 = lits[0], lits[1], lits[2], lits[3]
    p = lits[7]
    for err in [0.0, 0.3, 1.7, 12.0, 80.0]:
        s = err * a
        s = c3 * s * s * s + c1 * s - c2 * s * s
        want = 100.0 - 10.0 * s ** p if s > 0 else 100.0
        assert orc.remap(err) == want


def test_score_pools_through_the_pinned_pieces():
    """orc_ssimulacra2's score = remap(weighted_error(terms)) of its own terms."""
    rng = np.random.default_rng(4)
    ref = rng.random((3, 40, 56)).astype(F32)
    cand = np.clip(ref + rng.standard_normal(ref.shape).astype(F32) * F32(0.05), 0, 1).astype(F32)
    score, terms, raw = orc.ssimulacra2(ref, cand)
    ns = len(terms) // 18
    assert raw == rr.weighted_error(terms, ns)
    assert score == orc.remap(raw)


# ---------------------------------------------------------------------------
# Output depth thresholds (plan-render/lib.rs:153-216, :333-353)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("bits", [8, 10, 12, 16])
def test_depth_thresholds_are_the_level_boundaries(bits):
    """Every threshold t_q is the smallest f32 whose composed level
    (sRGB-encode, scale, round half away, clamp) reaches q, and every level's
    linear value is srgb_to_linear(q / max)."""
    thr, lut = orc.depth_tables(bits)
    levels = 1 << bits
    q = np.arange(1, levels)
    at = rr.composed_level(thr, bits)
    below = rr.composed_level(np.nextafter(thr, F32(-np.inf)), bits)
    assert np.all(at >= q)
    assert np.all(below < q)
    mx = F32(levels - 1)
    want = rr.srgb_to_linear32((np.arange(levels, dtype=F32) / mx).astype(F32))
    assert np.array_equal(lut.view(np.uint32), want.view(np.uint32))


# ---------------------------------------------------------------------------
# Hierarchical cover (lib.rs:4685-5302): brute force over every candidate
# ---------------------------------------------------------------------------

def _dct_fwd(block):
    return orc.dct2d(block[None])[0]


@pytest.mark.parametrize("wh,seed,gs", [((64, 64), 1, 900), ((64, 64), 2, 3000), ((96, 64), 3, 12000),
                                         ((61, 45), 4, 2000), ((40, 72), 5, 600)])
def test_cover_equals_the_brute_force_objective(wh, seed, gs):
    """The oracle's block map equals the quadtree's exact minimum of J with
    every candidate evaluated in full (merges strictly cheaper; calibrated
    rate model, quantizer-donor weights, AQ off)."""
    from jpxl_amd.synth import synthetic_rgb

    w, h = wh
    rgb = synthetic_rgb(w, h, seed, 8)
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    qp = orc.QuantParams(gs, 1, 4, 3, 3, 1, 4.0, 1.0)
    got = orc.cover(xyb, qp)
    want = rr.cover_brute_force(xyb, gs, 1, 3, 3, 4.0, _dct_fwd)
    assert np.array_equal(got, want)
    assert (got == 16).any() or (got == 32).any() or gs < 1000  # merges happen


@pytest.mark.parametrize("wh,seed,gs", [((64, 64), 6, 2000), ((61, 45), 7, 5000)])
def test_legacy_cover_equals_the_brute_force_objective(wh, seed, gs):
    """The fixed-quantizer request's cover objective (legacy rate proxy 1.0
    per bit with 2 / 34 / 34 fixed bits, flat cell weights, lambda scale 1)."""
    from jpxl_amd.synth import synthetic_rgb

    w, h = wh
    rgb = synthetic_rgb(w, h, seed, 8)
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    qp = orc.QuantParams(gs, 1, 16, 2, 2, 0, 1.0, 1.0)
    qp.cover_legacy = 1
    got = orc.cover(xyb, qp)
    want = rr.cover_brute_force(xyb, gs, 1, 2, 2, 1.0, _dct_fwd, calibrated=False, donor=False)
    assert np.array_equal(got, want)
    qp.cover_legacy = 0
    assert not np.array_equal(got, orc.cover(xyb, qp)) or gs > 4000  # the objectives differ
