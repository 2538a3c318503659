"""CPU tests: the oracle pinned against the reference's known-answer tests,
closed forms and invariants, plus the C ABI's exported symbols.

Each check cites the reference test or definition it pins (paths relative to
JPXL/crates/).  The planted-mistake tests at the end show that a plausible
slip in the restatement fails these checks.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle_binding as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# --- cube roots (jpxl-core/src/simd.rs:23-67, jpxl-perceptual/src/color.rs:46-61)

def test_metric_cbrt_kats():
    # jpxl-perceptual/src/color.rs:138-154
    for v in [1e-6, 0.0037930734, 0.1, 0.5, 1.0, 8.0, 123.456, 1e6]:
        v32 = float(np.float32(v))
        got = orc.cbrt_metric(v32)
        want = math.copysign(abs(v32) ** (1.0 / 3.0), v32)
        assert abs(got - want) <= want * 1e-6
        assert orc.cbrt_metric(-v32) == -got
    assert orc.cbrt_metric(0.0) == 0.0
    assert orc.cbrt_metric(27.0) == 3.0


def test_metric_cbrt_is_the_f32_rounding_of_the_f64_root():
    # Four Newton steps from the bit-hack estimate converge to the f64 cube
    # root, so the metric's f32 result is that root rounded once to f32 (the
    # device computes it by one Halley step from the f32 library root).
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.random(20000, dtype=np.float32) * 3.0 + 0.0037,
                         np.geomspace(1e-6, 1e6, 2001).astype(np.float32)])
    for x in xs:
        want = np.float32(np.cbrt(np.float64(x)))
        assert orc.cbrt_metric(float(x)) == float(want), x


def test_core_cbrt_is_correctly_rounded_on_a_sweep():
    # The musl algorithm (estimate + two Halley steps in f64) is correctly
    # rounded for normal inputs; compare with an exact f64 cube root rounded
    # once to f32 on a log sweep including subnormals.
    xs = np.geomspace(1e-40, 1e30, 4001).astype(np.float32)
    for x in xs:
        x = float(x)
        got = orc.cbrt_core(x)
        want = float(np.float32(np.cbrt(np.float64(x))))
        assert abs(got - want) <= abs(want) * 1.2e-7, x
    assert orc.cbrt_core(27.0) == 3.0
    assert orc.cbrt_core(-8.0) == -2.0
    assert orc.cbrt_core(0.0) == 0.0
    assert math.isinf(orc.cbrt_core(float("inf")))


# --- transfer function and XYB (jpxl-core/src/color.rs:820-1063)

def test_srgb_landmarks_and_round_trip():
    assert abs(orc.lib().orc_srgb_to_linear(0.5) - 0.21404114) < 1e-7
    assert abs(orc.lib().orc_linear_to_srgb(0.21404114) - 0.5) < 1e-6
    for v in range(256):
        lin = orc.lib().orc_srgb_to_linear(v / 255.0)
        back = orc.lib().orc_linear_to_srgb(lin)
        assert round(back * 255.0) == v
    # odd symmetry (NEGATIVES_TAKE_THE_LINEAR_SEGMENT[0] == false)
    assert orc.lib().orc_srgb_to_linear(-0.5) == -orc.lib().orc_srgb_to_linear(0.5)


def test_xyb_black_gray_and_inverse_round_trip():
    x, y, b = orc.xyb_pixel([0.0, 0.0, 0.0])
    assert abs(x) < 1e-6 and abs(y) < 1e-6 and abs(b) < 1e-6
    rng = np.random.default_rng(5)
    for g in [0.0, 0.01, 0.18, 0.5, 1.0]:
        x, y, b = orc.xyb_pixel([g, g, g])
        assert abs(x) < 1e-6  # neutral axis has X = 0
    for _ in range(200):
        rgb = rng.uniform(0, 1, 3).astype(np.float32)
        back = orc.opsin_inverse(orc.xyb_pixel(rgb))
        assert np.abs(back - rgb).max() < 1e-4


def test_positive_xyb_black_and_white():
    # jpxl-perceptual/src/color.rs:156-186
    x, y, b = orc.positive_xyb([0.0, 0.0, 0.0])
    assert abs(x - 0.42) < 1e-5 and abs(y - 0.01) < 1e-5 and abs(b - 0.55) < 1e-5
    w = orc.positive_xyb([1.0, 1.0, 1.0])
    assert ((w >= 0) & (w <= 1)).all()


def test_prepare_frame_grayscale_flag_and_luts():
    rgb = np.stack([np.arange(256, dtype=np.uint8)] * 3, -1).reshape(16, 16, 3)
    xyb, lin, gray = orc.prepare_frame(rgb, 8)
    assert gray
    assert np.abs(xyb[0]).max() < 1e-6
    lut = np.array([orc.lib().orc_srgb_to_linear(np.float32(v) / np.float32(255.0)) for v in range(256)], np.float32)
    assert np.array_equal(lin[0].ravel(), lut)
    rgb[0, 0, 0] ^= 1
    assert not orc.prepare_frame(rgb, 8)[2]
    # 16-bit entry point on 8-bit-exact samples agrees with the 8-bit one when
    # bits = 8 semantics are requested through the divisor 2^bits - 1.
    rgb16 = rgb.astype(np.uint16)
    xyb16, _, _ = orc.prepare_frame(rgb16, 16)
    assert np.isfinite(xyb16).all()


# --- DCT (jpxl-core/src/dct.rs:1375-1404, :1572, :1689, :1940-1972, :2040)

def naive2d(block, inverse=False):
    n = block.shape[0]
    m = np.array([[math.sqrt((1 if u == 0 else 2) / n) * math.cos(math.pi * (2 * x + 1) * u / (2 * n))
                   for x in range(n)] for u in range(n)])
    b = block.astype(np.float64)
    if not inverse:
        # column pass (vertical) then row pass, transposed orientation: out[k*n + r]
        t = (m @ b) / math.sqrt(n)          # t[u, c]
        out = (m @ t.T) / math.sqrt(n)      # out[k, u]  (k horizontal freq)
        return out
    t = (m.T @ b) * math.sqrt(n)            # inverse along first index
    return (m.T @ t.T) * math.sqrt(n)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_butterfly_matches_the_f64_definition(n):
    rng = np.random.default_rng(n)
    blocks = rng.uniform(-1, 1, (6, n, n)).astype(np.float32)
    got = orc.dct2d(blocks)
    for i in range(len(blocks)):
        want = naive2d(blocks[i])
        assert np.abs(got[i] - want).max() < 2e-6 * n
    back = orc.dct2d(got, inverse=True)
    assert np.abs(back - blocks).max() < 1e-5


@pytest.mark.parametrize("n", [8, 16, 32])
def test_dc_only_inverts_to_the_dc_value(n):
    c = np.zeros((1, n, n), np.float32)
    c[0, 0, 0] = 3.25
    s = orc.dct2d(c, inverse=True)[0]
    assert np.abs(s - 3.25).max() < 1e-5


@pytest.mark.parametrize("n", [8, 16, 32])
def test_parseval_with_normative_scale(n):
    rng = np.random.default_rng(7 * n)
    b = rng.uniform(-1, 1, (1, n, n)).astype(np.float32)
    c = orc.dct2d(b)[0].astype(np.float64)
    # normative = orthonormal / sqrt(n) per pass -> energy / n^2
    assert abs((c ** 2).sum() * n * n - (b.astype(np.float64) ** 2).sum()) < 1e-4 * n * n


def test_layout_first_row_holds_horizontal_frequencies():
    # dct_2d_in_place keeps the transposed orientation for squares: a pattern
    # that varies only across x has its energy in coefficient row 1 (k = 1).
    n = 8
    x = np.arange(n)
    b = np.tile(np.cos(np.pi * (2 * x + 1) / (2 * n)), (n, 1)).astype(np.float32)[None]
    c = orc.dct2d(b)[0]
    big = np.argwhere(np.abs(c) > 1e-3)
    assert big.tolist() == [[1, 0]]


def test_dct_iv_tables_are_orthonormal():
    for m in (4, 8, 16):
        t = orc.dct4_table(m).astype(np.float64)
        assert np.abs(t @ t.T - np.eye(m)).max() < 1e-6
        want = np.array([[math.sqrt(2 / m) * math.cos(math.pi * (2 * i + 1) * (2 * u + 1) / (4 * m))
                          for i in range(m)] for u in range(m)])
        assert np.abs(t - want).max() < 6e-8


def test_naive_1d_pair_round_trips():
    x = np.random.default_rng(1).uniform(-1, 1, 16)
    assert np.abs(orc.dct1d_naive(orc.dct1d_naive(x), True) - x).max() < 1e-12


def test_scale_f_closed_form():
    # jpxl-core/src/varblock.rs:789-811: ScaleF(0, n) = 1.
    for n in (1, 2, 4):
        assert orc.lib().orc_scale_f(0, n) == 1.0
    c, n = 1, 2
    b = 16.0
    inv = math.cos(c * math.pi / (2 * b)) * math.cos(c * math.pi / b) * math.cos(2 * c * math.pi / b)
    assert abs(orc.lib().orc_scale_f(c, n) - 1.0 / inv) < 1e-6


# --- dequantization and quantizer (jpxl-core/src/reconstruct.rs:122-163,
#     jpxl-encode-policy/src/quantize.rs:1753-2000)

def test_reconstruct_kats():
    o = orc.lib()
    o.orc_hf_multiplier.restype = ctypes.c_float
    o.orc_qm_multiplier.restype = ctypes.c_float
    o.orc_bias_adjust.restype = ctypes.c_float
    o.orc_bias_adjust.argtypes = [ctypes.c_int32, ctypes.c_float]
    assert o.orc_hf_multiplier(32768, 2) == 1.0
    assert o.orc_hf_multiplier(1, 1) == 65536.0
    assert o.orc_qm_multiplier(3) == np.float32(0.8)
    assert o.orc_qm_multiplier(2) == 1.0
    assert abs(o.orc_bias_adjust(1, 0.9) - 0.9) < 1e-7
    assert abs(o.orc_bias_adjust(10, 0.9) - (10 - 0.0145)) < 1e-6
    m = np.zeros(3, np.float32)
    o.orc_lf_multipliers(32768, 16, orc.P(m))
    want = [65536.0 * w / 128.0 / (32768 * 16) for w in (1 / 32, 1 / 4, 1 / 2)]
    assert np.allclose(m, want, rtol=1e-7)


@pytest.mark.parametrize("side", [8, 16, 32])
def test_dequant_matrices_shape(side):
    p0 = {8: (3150.0, 560.0, 512.0), 16: (8996.8725711814115328, 3191.48366296844234752, 1157.50408145487200256),
          32: (15718.40830982518931456, 7305.7636810695983104, 3803.53173721215041536)}[side]
    for c in range(3):
        m = orc.dequant_matrix(side, c)
        assert m[0, 0] == np.float32(1.0 / p0[c])  # bands[0] at the DC corner
        assert np.array_equal(m, m.T)             # radial: symmetric
        assert (m > 0).all()
    # Y params after the base are <= 0, so Mult < 1 and the weights fall with
    # distance: the dequantization step grows toward the highest frequency.
    m = orc.dequant_matrix(8, 1)
    assert m[0, 0] < m[7, 7]


def jpeg_zigzag(n):
    order = []
    for s in range(2 * n - 1):
        cells = [(y, s - y) for y in range(n) if 0 <= s - y < n]
        if s % 2 == 0:
            cells = cells[::-1]  # even diagonals run bottom-left to top-right
        order += [y * n + x for (y, x) in cells]
    return order


def test_natural_order_is_zigzag_for_dct8_and_llf_first():
    o8 = orc.natural_order(8)
    assert list(o8) == jpeg_zigzag(8)[:1] + [c for c in jpeg_zigzag(8) if c != 0]
    for side in (16, 32):
        o = orc.natural_order(side)
        assert sorted(o.tolist()) == list(range(side * side))
        n = side // 8
        assert sorted(o[: n * n].tolist()) == sorted(y * side + x for y in range(n) for x in range(n))


def test_donor_weights_mean_one():
    for side in (8, 16, 32):
        w = orc.donor_weights(side).astype(np.float64).reshape(side, side)
        n = side // 8
        mask = np.ones_like(w, bool)
        mask[:n, :n] = False
        assert abs(w[mask].mean() - 1.0) < 1e-6
        assert (w[~mask] == 1.0).all()


def test_hf_choose_nearest_rule():
    o = orc.lib()
    o.orc_hf_choose.restype = ctypes.c_int32
    o.orc_hf_choose.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int]
    o.orc_hf_recon.restype = ctypes.c_float
    o.orc_hf_recon.argtypes = [ctypes.c_int32, ctypes.c_float, ctypes.c_int]
    rng = np.random.default_rng(3)
    for _ in range(3000):
        step = float(np.float32(rng.uniform(0.001, 1.0)))
        t = float(np.float32(rng.uniform(-40, 40) * step))
        c = int(rng.integers(0, 3))
        q = o.orc_hf_choose(t, step, c)
        err = abs(o.orc_hf_recon(q, step, c) - np.float32(t))
        # never worse than naive rounding, nor than its neighbours
        for cand in (0, round(t / step), q - 1, q + 1):
            assert err <= abs(o.orc_hf_recon(cand, step, c) - np.float32(t)) + 1e-7 * abs(t)
    assert o.orc_hf_choose(0.0, 0.5, 1) == 0


def test_lf_quantizer_inverts_its_recon():
    xyb = np.zeros((3, 8, 8), np.float32)
    bmap = orc.fixed_dct8_map(8, 8)
    coeff = np.zeros((3, 8, 8), np.float32)
    coeff[:, 0, 0] = [0.01, 0.5, -0.2]
    q = orc.QuantParams(2000, 1, 4, 2, 2, 0, 4.0, 1.0)
    quant, lf = orc.quantize_frame(8, 8, bmap, coeff, q)
    assert (quant == 0).all()
    m = np.zeros(3, np.float32)
    orc.lib().orc_lf_multipliers(2000, 4, orc.P(m))
    for c in range(3):
        assert abs(lf[c, 0, 0] * m[c] - coeff[c, 0, 0]) <= 0.5 * m[c] + 1e-9 or c != 1


# --- restoration (jpxl-core/src/restoration.rs:791-869)

def test_gaborish_preserves_constants_and_kernel_sums_to_one():
    p = np.full((3, 9, 11), 0.37, np.float32)
    out = orc.gaborish(p)
    assert np.abs(out - 0.37).max() < 1e-6


@pytest.mark.parametrize("step", [0, 1, 2])
def test_epf_constant_and_skip(step):
    p = np.full((3, 16, 24), 0.25, np.float32)
    sigma = np.full((2, 3), 2.0, np.float32)
    assert np.abs(orc.epf_step(step, p, sigma) - 0.25).max() < 1e-7
    rng = np.random.default_rng(step)
    q = rng.uniform(0, 1, (3, 16, 24)).astype(np.float32)
    assert np.array_equal(orc.epf_step(step, q, np.zeros((2, 3), np.float32)), q)


def test_epf_weight_of_an_isolated_edge():
    # A hard vertical edge: pixels far from it are unchanged (every tap has
    # distance 0 within a flat region), pixels at it are pulled toward the
    # side whose taps are similar.
    p = np.zeros((3, 16, 16), np.float32)
    p[:, :, 8:] = 1.0
    sigma = np.full((2, 2), 5.0, np.float32)
    out = orc.epf_step(1, p, sigma)
    assert np.array_equal(out[:, :, :5], p[:, :, :5])
    assert np.array_equal(out[:, :, 11:], p[:, :, 11:])
    assert ((out >= 0) & (out <= 1)).all()


# --- SSIMULACRA2 (jpxl-perceptual/tests/metric.rs, src/blur.rs:685-795)

def fir_taps():
    # impulse response of the recursion, taken from the oracle's own blur
    e = np.zeros((1, 41), np.float32)
    e[0, 20] = 1.0
    line = orc.blur(np.repeat(e, 41, 0))[20]  # vertical pass of a column of ones... use row
    return line


def test_blur_equals_the_nine_tap_truncated_cosine_fir():
    # The recursion with poles at 2cos(k pi/10), k odd, is exactly a 9-tap FIR
    # in exact arithmetic: check the f64-state recursion against the direct
    # truncated-cosine sum on random planes (zero padded).
    n = 5
    taps = np.zeros(9)
    mul_in = [0.05529523572608661, -0.05883668702694998, 0.012955819110517063]
    for t in range(-4, 5):
        taps[t + 4] = sum(a * math.cos(t * k * math.pi / (2 * n)) * math.sin(k * math.pi / 2) / math.sin(k * math.pi / (2 * n))
                          for a, k in zip(mul_in, (1, 3, 5)))
    assert abs(taps.sum() - 1.0) < 1e-9
    rng = np.random.default_rng(11)
    a = rng.uniform(0, 1, (37, 53)).astype(np.float32)
    got = orc.blur(a).astype(np.float64)
    pad = np.pad(a.astype(np.float64), 4)
    hor = sum(taps[t] * pad[4:-4, t : t + 53] for t in range(9))
    hor = np.pad(hor, ((4, 4), (0, 0)))
    want = sum(taps[t] * hor[t : t + 37, :] for t in range(9))
    assert np.abs(got - want).max() < 2e-6


def _exact_split_lines(u, carry=True):
    """The device's form of the blur recursion over the rows of u (f32):
    y = f32(F + E) with F the 9-tap blur in f64 and E the alternating
    stride-10 prefix of Phi, Phi the resonators' 10-tap response to the
    rounding errors of the f32 pair sums (exact by TwoSum)."""
    mul_in = [0.05529523572608661, -0.05883668702694998, 0.012955819110517063]
    mul_prev = [1.902113032590307, 1.1755705045849463, 1.2246467991473532e-16]
    g = []
    p1, p2 = [0.0] * 3, [0.0] * 3
    for r in range(10):
        o = [(1.0 if r == 0 else 0.0) * mul_in[k] + mul_prev[k] * p1[k] - p2[k] for k in range(3)]
        p2, p1 = p1, o
        g.append(o[0] + o[1] + o[2])
    g32 = np.array(g, np.float32)
    n, L = u.shape
    pad = np.zeros((n, L + 24), np.float32)
    pad[:, 12:12 + L] = u
    at = lambda i: pad[:, 12 + i]
    eps = np.zeros((n, L + 4), np.float32)
    for p in range(L + 4):
        a, b = at(p - 10), at(p)
        t = a + b
        bv = t - a
        eps[:, p] = -((a - (t - bv)) + (b - bv))
    E = np.zeros((n, L))
    out = np.zeros((n, L), np.float32)
    for m in range(L):
        phi = np.zeros(n, np.float32)
        for r in range(9, -1, -1):
            if m + 4 - r >= 0:
                phi = (g32[r] * eps[:, m + 4 - r] + phi).astype(np.float32)
        E[:, m] = phi.astype(np.float64) - (E[:, m - 10] if m >= 10 else 0.0)
        F = g[4] * at(m).astype(np.float64)
        for t in range(1, 5):
            F = F + g[4 + t] * (at(m - t).astype(np.float64) + at(m + t).astype(np.float64))
        out[:, m] = (F + (E[:, m] if carry else 0.0)).astype(np.float32)
    return out


def test_blur_exact_split_equals_the_recursion():
    # The form the device evaluates (an exact blur plus the carried pair
    # roundings) reproduces the recursion on long lines, where the exact
    # 9-tap blur alone drifts from it.  It rounds the recursion's exact value;
    # the reference's f64 resonators land on the other f32 neighbour for
    # about one output in 10^6 (none on this plane).
    rng = np.random.default_rng(12)
    h, w = 24, 700
    yy, xx = np.mgrid[0:h, 0:w]
    a = (0.45 + 0.2 * np.sin(xx / 17.0) * np.cos(yy / 5.0) + rng.random((h, w)) * 0.05).astype(np.float32)
    for plane in (a, (a * a).astype(np.float32)):
        want = orc.blur(plane)
        hor = _exact_split_lines(plane)
        got = _exact_split_lines(np.ascontiguousarray(hor.T)).T
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the carried roundings matter: the blur without them differs
    hor = _exact_split_lines(a, carry=False)
    got = _exact_split_lines(np.ascontiguousarray(hor.T), carry=False).T
    assert not np.array_equal(got, orc.blur(a))


def test_blur_constant_interior_and_dark_border():
    out = orc.blur(np.ones((48, 64), np.float32))
    assert abs(out[24, 32] - 1.0) < 2e-3
    assert out[0, 0] < out[24, 32]


def test_product_blur_matches_materialised():
    rng = np.random.default_rng(2)
    a = rng.uniform(0, 1, (51, 73)).astype(np.float32)
    b = rng.uniform(0, 1, (51, 73)).astype(np.float32)
    assert np.array_equal(orc.blur(a, b), orc.blur(a * b))


def _noise_img(h, w, seed):
    return np.random.default_rng(seed).uniform(0, 1, (3, h, w)).astype(np.float32)


def test_identical_images_score_exactly_100():
    ref = _noise_img(64, 80, 1)
    s, terms, raw = orc.ssimulacra2(ref, ref)
    assert s == 100.0 and raw == 0.0 and not terms.any()


def test_scale_counts():
    ref = _noise_img(192, 256, 2)
    assert len(orc.ssimulacra2(ref, ref)[1]) == 18 * 6
    small = _noise_img(8, 8, 3)
    assert len(orc.ssimulacra2(small, small)[1]) == 18 * 2


def test_score_falls_with_noise():
    ref = _noise_img(64, 64, 4) * 0.5 + 0.25
    rng = np.random.default_rng(9)
    last = 100.0
    for amp in (0.01, 0.03, 0.1):
        cand = np.clip(ref + rng.uniform(-amp, amp, ref.shape).astype(np.float32), 0, 1)
        s = orc.ssimulacra2(ref, cand)[0]
        assert s < last
        last = s


def test_remap_landmarks():
    assert orc.remap(0.0) == 100.0
    assert orc.remap(1.0) < orc.remap(0.5) < 100.0


# --- planted mistakes: a plausible slip must fail the checks above -----------

def test_a_transposed_coefficient_layout_would_fail_the_layout_check():
    n = 8
    x = np.arange(n)
    b = np.tile(np.cos(np.pi * (2 * x + 1) / (2 * n)), (n, 1)).astype(np.float32)[None]
    c = orc.dct2d(b)[0].T  # the plausible mistake
    assert np.argwhere(np.abs(c) > 1e-3).tolist() != [[1, 0]]


def test_a_dropped_bias_term_would_fail_the_black_point():
    x, y, b = orc.positive_xyb([0.0, 0.0, 0.0])
    assert abs((b - 0.55) - (-0.55)) > 0.5  # without +0.55 the check above fails


# --- the C ABI exports every declared symbol ---------------------------------

def test_library_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "jpxlg.h")).read()
    names = set(re.findall(r"\b(jpxlg_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 18
    so = os.path.join(ROOT, "jpxl_amd", "libjpxlg.so")
    if not os.path.exists(so):
        pytest.skip("libjpxlg.so not built")
    lib = ctypes.CDLL(so)
    for n in sorted(names):
        assert hasattr(lib, n), n
    from jpxl_amd import _lib
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_f64_frame_dct_matches_the_butterfly_oracle():
    # The precision study's float64 closed form (jpxl_amd.precision) and the
    # oracle's f32 butterflies agree to f32 rounding on mixed block maps, so
    # the closed form has the reference's scale, transposition and footprint.
    from jpxl_amd.precision import error_report, forward_frame_f64
    from jpxl_amd.synth import synthetic_rgb

    for w, h in [(64, 64), (61, 37), (130, 70)]:
        rgb = synthetic_rgb(w, h, 5, 8)
        xyb, _, _ = orc.prepare_frame(rgb, 8)
        bmap = orc.mixed_map(w, h, seed=w)
        rep = error_report(orc.forward_frame(xyb, bmap), forward_frame_f64(xyb, bmap), bmap)
        assert rep and all(r["max_abs"] < 5e-7 for r in rep.values()), rep


# --- encoder-side Gaborish sharpen (jpxl-encode/src/vardct/gaborish.rs:66-315)

def _plain_conv3(plane):
    # the normalised J.1 kernel applied with mirrored edges, in float64
    w1, w2 = 0.115169525, 0.061248592
    k = np.array([[w2, w1, w2], [w1, 1.0, w1], [w2, w1, w2]]) / (1 + 4 * w1 + 4 * w2)
    h, w = plane.shape
    ys = np.array([[abs(y + d) if y + d >= 0 else -(y + d) - 1 for d in (-1, 0, 1)] for y in range(h)])
    ys = np.where(ys >= h, 2 * h - 1 - ys, ys)
    xs = np.array([[abs(x + d) if x + d >= 0 else -(x + d) - 1 for d in (-1, 0, 1)] for x in range(w)])
    xs = np.where(xs >= w, 2 * w - 1 - xs, xs)
    p = plane.astype(np.float64)
    out = np.zeros_like(p)
    for a in range(3):
        for b in range(3):
            out += k[a, b] * p[ys[:, a]][:, xs[:, b]]
    return out


@pytest.mark.parametrize("wh", [(13, 9), (64, 40), (5, 7)])
def test_encoder_gaborish_forward_is_the_normalised_kernel(wh):
    # Every branch (mirrored border, four-wide interior groups, scalar tail)
    # evaluates the same normalised 3x3 kernel up to f32 rounding.
    w, h = wh
    rng = np.random.default_rng(w * 7 + h)
    planes = rng.uniform(-1, 1, (1, h, w)).astype(np.float32)
    got = orc.gaborish_forward_enc(planes)[0]
    assert np.abs(got - _plain_conv3(planes[0])).max() < 2e-6


def test_gaborish_sharpen_inverts_the_forward_filter():
    # Jacobi on K x = f: the residual |K x - f| shrinks with the step count
    # and the 12-step result is close to an exact inverse; a constant plane
    # (K preserves constants) is a fixed point.
    rng = np.random.default_rng(3)
    base = rng.uniform(0, 1, (1, 48, 61)).astype(np.float32)
    f = orc.gaborish_forward_enc(base)
    res = []
    for it in (1, 4, 12):
        x = orc.gaborish_sharpen(f, it)
        res.append(np.abs(orc.gaborish_forward_enc(x) - f).max())
    assert res[0] > res[1] > res[2] and res[2] < 1e-3
    assert np.abs(orc.gaborish_sharpen(f, 12) - base).max() < 5e-3
    const = np.full((1, 17, 23), 0.375, np.float32)
    assert np.abs(orc.gaborish_sharpen(const, 12) - 0.375).max() < 1e-6


# --- adaptive quantization (jpxl-encode-policy/src/analysis.rs, field.rs,
# lib.rs:1847-1944)

def _half_flat_half_noise(w: int, h: int) -> np.ndarray:
    """field.rs:441-456: left half gray 128, right half hashed noise."""
    x = np.arange(w, dtype=np.uint64)[None, :]
    y = np.arange(h, dtype=np.uint64)[:, None]
    m32 = np.uint64(0xFFFFFFFF)
    hsh = (((x * np.uint64(0x9E37)) & m32) + ((y * np.uint64(0x79B9)) & m32)) & m32
    hsh = (hsh * np.uint64(0x85EBCA6B)) & m32
    noise = (96 + ((hsh >> np.uint64(24)) & np.uint64(0x3F))).astype(np.uint8)
    luma = np.where(x < w // 2, np.uint8(128), noise).astype(np.uint8)
    return np.repeat(luma[:, :, None], 3, axis=2)


def _xyb_of_rgb8(rgb: np.ndarray) -> np.ndarray:
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    return xyb


def _xyb_of_linear(planes: np.ndarray) -> np.ndarray:
    _, h, w = planes.shape
    out = np.zeros((3, h, w), np.float32)
    for y in range(h):
        for x in range(w):
            out[:, y, x] = orc.xyb_pixel(planes[:, y, x])
    return out


def test_atlas_flat_frame_has_zero_variance():
    # analysis.rs:493-518: 20x12 flat linear 0.25 -> a 3x2 grid, all zero
    xyb = _xyb_of_linear(np.full((3, 12, 20), 0.25, np.float32))
    var = orc.atlas_variance(xyb)
    assert var.shape == (3, 2, 3)
    assert np.all(np.abs(var) < 1e-9)


def test_atlas_edge_variance_stays_in_its_atom():
    # analysis.rs:521-542: a step at x = 4 lies in atom column 0 only
    plane = np.full((8, 16), 0.1, np.float32)
    plane[:, 4:8] = 0.9
    var = orc.atlas_variance(_xyb_of_linear(np.stack([plane] * 3)))
    assert var[1, 0, 0] > var[1, 0, 1]
    assert abs(var[1, 0, 1]) < 1e-9


def test_atlas_variance_matches_the_population_variance():
    # closed form: E[x^2] - E[x]^2 over the samples that exist (partial edge
    # atoms included), f64, narrowed to f32
    rng = np.random.default_rng(11)
    xyb = rng.random((3, 21, 27), dtype=np.float32)
    var = orc.atlas_variance(xyb)
    for c in range(3):
        for ay in range(3):
            for ax in range(4):
                blk = xyb[c, 8 * ay:8 * ay + 8, 8 * ax:8 * ax + 8].astype(np.float64)
                want = max(float(np.mean(blk * blk) - np.mean(blk) ** 2), 0.0)
                assert abs(float(var[c, ay, ax]) - want) <= 1e-7 * max(want, 1e-3)


def test_masking_points_coarse_at_noise_and_fine_at_flat():
    # field.rs:459-464
    adj, active, fine = orc.aq_field(_xyb_of_rgb8(_half_flat_half_noise(128, 64)), orc.AQ_MASKING)
    assert active and not fine
    assert adj[1, 1] < 0.0
    assert adj[1, 12] > 0.0


def test_uniform_is_the_exact_mirror_of_masking():
    # field.rs:467-478
    xyb = _xyb_of_rgb8(_half_flat_half_noise(128, 64))
    m, _, _ = orc.aq_field(xyb, orc.AQ_MASKING)
    u, _, _ = orc.aq_field(xyb, orc.AQ_UNIFORM)
    for x, y in [(1, 1), (12, 1), (6, 4)]:
        assert abs(m[y, x] + u[y, x]) < 1e-6


def test_off_and_flat_frames_give_no_field():
    # field.rs:481-493: Off has no field; a flat frame's field is neutral
    # (every multiplier the baseline), which AqSetup treats as off
    xyb = _xyb_of_rgb8(np.full((64, 64, 3), 100, np.uint8))
    _, active, _ = orc.aq_field(xyb, orc.AQ_OFF)
    assert not active
    adj, active, _ = orc.aq_field(xyb, orc.AQ_MASKING)
    assert not active
    for ax, ay in [(0, 0), (3, 3), (7, 7)]:
        assert orc.aq_footprint_mul(adj, ax, ay, 1, 1, 2, False, False) == 2


def test_factorized_multipliers_and_refine_only_cap():
    # field.rs:496-515
    adj, _, _ = orc.aq_field(_xyb_of_rgb8(_half_flat_half_noise(128, 64)), orc.AQ_MASKING)
    fine = orc.aq_footprint_mul(adj, 1, 1, 1, 1, 2, False, False)
    coarse = orc.aq_footprint_mul(adj, 12, 1, 1, 1, 2, False, False)
    assert 2 < fine <= 4
    assert coarse < 2
    assert orc.aq_footprint_mul(adj, 12, 1, 1, 1, 2, True, False) == 2
    assert orc.aq_footprint_mul(adj, 1, 1, 1, 1, 2, True, False) == fine


def test_fine_lattice_and_edge_refine_never_coarsens():
    # field.rs:609-657 (the edge arm at the default tuning)
    xyb = _xyb_of_rgb8(_half_flat_half_noise(128, 64))
    adj, active, fine = orc.aq_field(xyb, orc.AQ_FINE_MASKING)
    assert active and fine
    muls = {orc.aq_footprint_mul(adj, x, y, 1, 1, 16, False, True) for y in range(8) for x in range(16)}
    assert min(muls) >= 8 and max(muls) <= 32
    assert orc.aq_footprint_mul(adj, 1, 1, 1, 1, 16, False, True) > orc.aq_footprint_mul(adj, 12, 1, 1, 1, 16,
                                                                                           False, True)
    edge, _, efine = orc.aq_field(xyb, orc.AQ_EDGE_REFINE)
    assert efine
    for y in range(8):
        for x in range(16):
            assert orc.aq_footprint_mul(edge, x, y, 1, 1, 16, False, True) >= 16
    assert orc.aq_footprint_mul(edge, 1, 3, 1, 1, 16, False, True) == 16


def test_mul_of_closed_forms():
    # field.rs:383-395: round(baseline * 2^-adj), clamped to [1, 2 baseline]
    assert [orc.aq_mul_of(2, a) for a in (-1.0, -0.5, 0.0, 0.5, 1.0)] == [4, 3, 2, 1, 1]
    for a in np.linspace(-1.0, 1.0, 33, dtype=np.float32):
        want = min(max(int(math.floor(16 * 2.0 ** (-float(a)) + 0.5)), 1), 32)
        assert orc.aq_mul_of(16, float(a)) == want
    assert orc.aq_mul_of(1, 3.0) == 1  # never below 1
    assert orc.aq_mul_of(5, -4.0) == 10  # capped one octave finer


def test_footprint_mean_snaps_to_half_octaves():
    # field.rs:338-378: f32 mean in row-major order, half-octave snap (ties
    # away from zero), zero outside the grid
    rng = np.random.default_rng(3)
    adj = (rng.random((9, 13), dtype=np.float32) * 2 - 1).astype(np.float32)
    for (bx, by, n) in [(0, 0, 1), (3, 2, 2), (8, 4, 4), (12, 8, 2), (11, 7, 4)]:
        s = np.float32(0.0)
        for dy in range(n):
            for dx in range(n):
                x, y = bx + dx, by + dy
                s = np.float32(s + (adj[y, x] if x < 13 and y < 9 else np.float32(0.0)))
        a = np.float32(s / np.float32(n * n))
        snapped = math.copysign(math.floor(abs(float(a * np.float32(2.0))) + 0.5), float(a)) / 2.0
        assert orc.aq_footprint_mul(adj, bx, by, n, n, 2, False, False) == orc.aq_mul_of(2, snapped)
        assert orc.aq_footprint_mul(adj, bx, by, n, n, 2, True, False) == orc.aq_mul_of(2, min(snapped, 0.0))


def test_aq_setup_factorization():
    # lib.rs:1859-1908
    assert orc.aq_setup(32768, 16, 1, False) == (16384, 32, 2, False)
    assert orc.aq_setup(32769, 16, 1, False) == (16384, 32, 2, False)  # snapped down
    assert orc.aq_setup(1, 16, 1, False) == (1, 16, 1, True)  # snapped < factor
    assert orc.aq_setup(4096, 40000, 1, False) == (4096, 40000, 1, True)  # quant_lf x2 not representable
    assert orc.aq_setup(32768, 4, 3, True) == (2048, 64, 48, False)
    assert orc.aq_setup(17, 4, 3, True) == (1, 64, 48, False)
    assert orc.aq_setup(15, 4, 3, True) == (15, 4, 3, True)


def test_constant_aq_field_equals_a_global_multiplier():
    # With every varblock at one multiplier m (a constant field) and the
    # nearest rule (no Lagrange weights), quantize and render equal the
    # plain path at hf_mul = m: the per-varblock multiplier is the only
    # thing AQ changes.
    w, h = 72, 56
    rng = np.random.default_rng(9)
    rgb = (rng.random((h, w, 3)) * 255).astype(np.uint8)
    xyb = _xyb_of_rgb8(rgb)
    bmap = orc.mixed_map(w, h, 4)
    coeff = orc.forward_frame(xyb, bmap)
    bh, bw = bmap.shape
    for c, m in [(-1.0, 4), (-0.5, 3), (1.0, 1)]:
        adj = np.full((bh, bw), c, np.float32)
        q_aq = orc.with_aq(orc.QuantParams(4096, 2, 32, 2, 2, 0, 1.0, 0.0), adj)
        q_pl = orc.QuantParams(4096, m, 32, 2, 2, 0, 1.0, 0.0)
        assert np.all(orc.aq_block_muls(w, h, bmap, q_aq)[bmap > 0] == m)
        qa, la = orc.quantize_frame(w, h, bmap, coeff, q_aq)
        qp, lp = orc.quantize_frame(w, h, bmap, coeff, q_pl)
        assert np.array_equal(qa, qp) and np.array_equal(la, lp)
        r_aq = orc.with_aq(orc.RenderParams(4096, 2, 32, 2, 2, 1, 0, 7, 8), adj)
        r_pl = orc.RenderParams(4096, m, 32, 2, 2, 1, 0, 7, 8)
        la_, xa = orc.render(w, h, bmap, qa, la, r_aq)
        lp_, xp = orc.render(w, h, bmap, qp, lp, r_pl)
        assert np.array_equal(la_, lp_) and np.array_equal(xa, xp)


def test_chroma_texture_alone_drives_the_masking_field():
    # field.rs:275-283: activity = luma + 0.35 * chroma.  With flat Y and
    # textured X/B only the chroma term sees texture, so a dropped chroma
    # term (a plausible slip) would leave this field neutral.
    rng = np.random.default_rng(2)
    xyb = np.zeros((3, 64, 64), np.float32)
    xyb[1] = 0.5
    xyb[0, :, 32:] = rng.random((64, 32), dtype=np.float32) * 0.02
    xyb[2, :, 32:] = 0.4 + rng.random((64, 32), dtype=np.float32) * 0.1
    xyb[2, :, :32] = 0.45
    adj, active, _ = orc.aq_field(xyb, orc.AQ_MASKING)
    assert active
    assert np.all(adj[:, 5:] > 0) and np.all(adj[:, :3] < 0)
