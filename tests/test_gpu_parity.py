"""GPU parity tests: every stage of the HIP path against the CPU oracle.

Integer and index outputs (quantized coefficients, LF planes) must be
bit-identical; float planes of the pixel path are compared bit for bit as well
(the device restates the reference's operation order exactly).  The GPU
blur evaluates the reference's recursion (its f32 pair sums and their
rounding included) exactly before the one f32 rounding of each output; the
reference's f64 resonators round a few ulps of f64 away from that value, which
turns about one output in 10^6 to the other f32 neighbour (the vertical pass
then carries that input change down its column).  SSIMULACRA2 scores of the
same planes therefore agree within PAIR_TOL; scores of rendered candidates
within SCORE_TOL, which also covers the rare one-ulp differences of the
device's positive-XYB cube root.
"""
import numpy as np
import pytest

import oracle_binding as orc

pytestmark = pytest.mark.gpu

SCORE_TOL = 2e-4
PAIR_TOL = 1e-7


def _terms_close(terms, oterms, tol=1e-6):
    # Terms of rendered candidates carry the cube root's rare one-ulp
    # differences (amplified by the 4-norms of near-zero maps), hence an
    # absolute bound.
    assert len(terms) == len(oterms)
    assert np.abs(terms - oterms).max() < tol


@pytest.fixture(scope="module")
def ctx():
    import jpxl_amd

    c = jpxl_amd.Context(0)
    yield c
    c.close()


def _rgb(w, h, seed, bits=8):
    from jpxl_amd.synth import synthetic_rgb

    return synthetic_rgb(w, h, seed, bits)


def _quantizer(gs=2000, mul=1, qlf=4, xqm=3, bqm=3, truncate=1):
    from jpxl_amd import Quantizer

    return Quantizer(gs, mul, qlf, xqm, bqm, truncate, 4.0, 1.0)


def _qp(q):
    return orc.QuantParams(q.global_scale, q.hf_mul, q.quant_lf, q.x_qm_scale, q.b_qm_scale, q.truncate,
                           q.lambda_scale, q.zero_token_bits)


SIZES = [(64, 64), (61, 37), (8, 8), (130, 97)]


@pytest.mark.parametrize("bits", [8, 12, 16])
@pytest.mark.parametrize("wh", SIZES)
def test_xyb_planes_bit_exact(ctx, wh, bits):
    import jpxl_amd
    from jpxl_amd import _lib

    w, h = wh
    rgb = _rgb(w, h, 3, bits)
    img = jpxl_amd.Image(ctx, rgb, bits)
    got = img.download(_lib.BUF_XYB)
    want, _, gray = orc.prepare_frame(rgb, bits)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert img.grayscale == gray


def test_grayscale_flag(ctx):
    import jpxl_amd

    g = np.repeat(_rgb(32, 24, 4)[..., :1], 3, axis=2)
    assert jpxl_amd.Image(ctx, g, 8).grayscale
    assert not jpxl_amd.Image(ctx, _rgb(32, 24, 4), 8).grayscale


@pytest.mark.parametrize("n", [8, 16, 32])
@pytest.mark.parametrize("inverse", [False, True])
def test_dct_blocks_bit_exact(ctx, n, inverse):
    rng = np.random.default_rng(n + 100 * inverse)
    blocks = rng.uniform(-1, 1, (37, n, n)).astype(np.float32)
    got = ctx.dct_blocks(blocks, inverse)
    want = orc.dct2d(blocks, inverse)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _image_and_oracle(ctx, w, h, seed, bits, bmap, cfl=None, lf_factors=(0, 0)):
    import jpxl_amd

    rgb = _rgb(w, h, seed, bits)
    img = jpxl_amd.Image(ctx, rgb, bits)
    cx, cb = (None, None) if cfl is None else cfl
    img.set_structure(bmap, cx, cb, lf_factors)
    xyb, lin, _ = orc.prepare_frame(rgb, bits)
    return img, xyb, lin


@pytest.mark.parametrize("wh", [(64, 64), (61, 37), (257, 131)])
def test_forward_coefficients_bit_exact(ctx, wh):
    from jpxl_amd import _lib

    w, h = wh
    bmap = orc.mixed_map(w, h, seed=w)
    img, xyb, _ = _image_and_oracle(ctx, w, h, 5, 8, bmap)
    got = img.download(_lib.BUF_COEFF)
    want = orc.forward_frame(xyb, bmap)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("form,tol", [(1, 2e-6), (2, 1e-4)])
def test_matrix_dct_forms_against_f64(ctx, form, tol):
    """The matrix-core DCT forms (FP32 products, bf16 hi/lo split) against the
    float64 closed form; switching back to the butterflies is bit-exact."""
    from jpxl_amd import _lib
    from jpxl_amd.precision import error_report, forward_frame_f64

    w, h = 257, 131
    bmap = orc.mixed_map(w, h, seed=7)
    img, xyb, _ = _image_and_oracle(ctx, w, h, 5, 8, bmap)
    want64 = forward_frame_f64(xyb, bmap)
    exact = img.download(_lib.BUF_COEFF)
    img.set_dct_form(form)
    got = img.download(_lib.BUF_COEFF)
    rep = error_report(got, want64, bmap)
    assert set(rep) == {"8", "16", "32"}
    assert all(r["max_abs"] < tol for r in rep.values()), rep
    assert not np.array_equal(got, exact) or form == 1
    img.set_dct_form(_lib.DCT_EXACT)
    assert np.array_equal(img.download(_lib.BUF_COEFF).view(np.uint32), exact.view(np.uint32))


@pytest.mark.parametrize("form", [1, 2])
def test_matrix_idct_render_close_to_exact(ctx, form):
    """Render through the matrix-core inverse: close to the bit-exact render."""
    from jpxl_amd import Restoration, _lib

    w, h = 200, 120
    bmap = orc.mixed_map(w, h, seed=3)
    img, xyb, _ = _image_and_oracle(ctx, w, h, 9, 8, bmap)
    img.quantize(_quantizer(gs=8000))
    img.render_score(Restoration(0, 0, 7))
    exact = img.download(_lib.BUF_RENDER_XYB)
    img.set_dct_form(form)
    img.quantize(_quantizer(gs=8000))
    img.render_score(Restoration(0, 0, 7))
    got = img.download(_lib.BUF_RENDER_XYB)
    assert np.abs(got - exact).max() < (1e-4 if form == 2 else 1e-5)


def _cfl(bmap, seed):
    th, tw = (bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8
    rng = np.random.default_rng(seed)
    return rng.integers(-20, 21, (th, tw)).astype(np.int8), rng.integers(-20, 21, (th, tw)).astype(np.int8)


QUANT_CASES = [
    dict(gs=2000, mul=1, qlf=4, xqm=3, bqm=3, truncate=1),
    dict(gs=2000, mul=1, qlf=4, xqm=2, bqm=2, truncate=0),
    dict(gs=300, mul=1, qlf=16, xqm=3, bqm=3, truncate=1),
    dict(gs=32768, mul=1, qlf=16, xqm=2, bqm=2, truncate=0),
    dict(gs=73728, mul=7, qlf=16, xqm=3, bqm=4, truncate=1),
]


@pytest.mark.parametrize("case", QUANT_CASES)
@pytest.mark.parametrize("mixed", [False, True])
def test_quantize_bit_exact(ctx, case, mixed):
    w, h = 203, 141
    bmap = orc.mixed_map(w, h, seed=7) if mixed else orc.fixed_dct8_map(w, h)
    cfl = _cfl(bmap, 11) if mixed else None
    lff = (5, -9) if mixed else (0, 0)
    img, xyb, _ = _image_and_oracle(ctx, w, h, 6, 8, bmap, cfl, lff)
    q = _quantizer(**case)
    img.quantize(q)
    plan = img.plan_download()
    coeff = orc.forward_frame(xyb, bmap)
    cx, cb = cfl if cfl else (None, None)
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, _qp(q), cx, cb, lff)
    assert np.array_equal(plan.lf, olf)
    assert np.array_equal(plan.quant, oq)
    assert np.array_equal(plan.block_map, bmap)


@pytest.mark.parametrize("wh,gs", [((256, 192), 1200), ((131, 77), 400), ((200, 136), 5000)])
def test_cover_block_map_matches_the_oracle(ctx, wh, gs):
    import jpxl_amd
    from jpxl_amd import _lib

    w, h = wh
    rgb = _rgb(w, h, 17)
    img = jpxl_amd.Image(ctx, rgb, 8)
    q = _quantizer(gs=gs)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL, False)
    plan_map = img.plan_download_map()
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    want = orc.cover(xyb, _qp(q))
    assert np.array_equal(plan_map, want)
    assert set(np.unique(want)) >= {0, 8}  # the synthetic image merges some blocks


@pytest.mark.parametrize("wh,gs", [((256, 192), 1200), ((131, 77), 5000)])
def test_legacy_cover_block_map_matches_the_oracle(ctx, wh, gs):
    """The fixed-quantizer request's cover objective (legacy rate proxy, flat
    cell weights, the nearest rule, lambda scale 1)."""
    import jpxl_amd
    from jpxl_amd import Quantizer, _lib

    w, h = wh
    rgb = _rgb(w, h, 18)
    img = jpxl_amd.Image(ctx, rgb, 8)
    q = Quantizer(gs, 1, 16, 2, 2, 0, 1.0, 1.0)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL_LEGACY, False)
    plan_map = img.plan_download_map()
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    qp = _qp(q)
    qp.cover_legacy = 1
    assert np.array_equal(plan_map, orc.cover(xyb, qp))
    img.close()


@pytest.mark.parametrize("effort,target,seed", [("balanced", 85.0, 31), ("fast", 70.0, 32), ("balanced", 60.0, 33),
                                                ("fast", 90.0, 34)])
def test_encode_quality_leaves_the_chosen_plan_on_its_structure(ctx, effort, target, seed):
    """Whatever the navigator did (an anchor at the predicted rung, a finalist
    rebuilt with its own structure), the plan left on the device is the
    oracle's quantization of the chosen rung over the structure the chosen
    probe was scored with, and rescoring it reproduces the reported score.
    Fast plans carry neutral CfL (planning_tiers, quality.rs:1739)."""
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    w, h = 200, 168
    rgb = _rgb(w, h, seed)
    img = jpxl_amd.Image(ctx, rgb, 8)
    res = img.encode_quality(target, effort)
    plan = img.plan_download()
    pol_qm = 2 if effort == "fast" else 3
    trunc = 0 if effort == "fast" else 1
    q = Quantizer(res.global_scale, res.hf_mul, res.quant_lf, pol_qm, pol_qm, trunc, 4.0, 1.0)
    # the structure the chosen probe was scored with
    srung = res.rung if res.structure_rebuilt else res.predicted_rung
    assert res.probe_fresh[0] == 1 and res.probe_rungs[0] == res.predicted_rung
    ref = jpxl_amd.Image(ctx, rgb, 8)
    sq = jpxl_amd.rung_quantizer(srung, effort)
    ref.plan_structure(sq, _lib.COVER_FIXED_DCT8 if effort == "fast" else _lib.COVER_HIERARCHICAL, effort != "fast")
    ref.quantize(q)
    want = ref.plan_download()
    assert np.array_equal(plan.block_map, want.block_map)
    assert np.array_equal(plan.cfl_x, want.cfl_x) and np.array_equal(plan.cfl_b, want.cfl_b)
    assert plan.lf_factors == want.lf_factors
    if effort == "fast":
        assert not plan.cfl_x.any() and not plan.cfl_b.any() and plan.lf_factors == (0, 0)
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    coeff = orc.forward_frame(xyb, plan.block_map)
    oq, olf = orc.quantize_frame(w, h, plan.block_map, coeff, _qp(q), plan.cfl_x, plan.cfl_b, plan.lf_factors)
    assert np.array_equal(plan.quant, oq) and np.array_equal(plan.lf, olf)
    assert img.render_score(Restoration(1, 0, 7))[0] == res.score
    img.close()
    ref.close()


def _flat_and_busy(w, h, seed):
    # left half one grey level (every candidate ties on rate alone), right
    # half synthetic detail
    rgb = _rgb(w, h, seed)
    rgb[:, : w // 2] = 128
    return rgb


@pytest.mark.parametrize("kind", ["synthetic", "flat_and_busy"])
@pytest.mark.parametrize("wh,gs", [((512, 384), 1500), ((257, 131), 700)])
def test_certified_cover_equals_the_exact_pass(ctx, wh, gs, kind, monkeypatch):
    # The certified double-double costs decide most regions; the exact pass
    # (the reference's sequential sums) redoes the rest.  Forcing every
    # region through the exact pass must give the same block map.
    import jpxl_amd
    from jpxl_amd import _lib

    w, h = wh
    rgb = _flat_and_busy(w, h, 5) if kind == "flat_and_busy" else _rgb(w, h, 5)
    q = _quantizer(gs=gs)
    img = jpxl_amd.Image(ctx, rgb, 8)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL, False)
    fast = img.plan_download_map()
    monkeypatch.setenv("JPXLG_COVER_ALL_EXACT", "1")
    ex_ctx = jpxl_amd.Context(0)
    try:
        img2 = jpxl_amd.Image(ex_ctx, rgb, 8)
        img2.plan_structure(q, _lib.COVER_HIERARCHICAL, False)
        exact = img2.plan_download_map()
    finally:
        ex_ctx.close()
    assert np.array_equal(fast, exact)
    if w <= 260:
        xyb, _, _ = orc.prepare_frame(rgb, 8)
        assert np.array_equal(fast, orc.cover(xyb, _qp(q)))


def _tinted(w, h, seed):
    g = _rgb(w, h, seed)[..., 1].astype(np.float32)
    return np.stack([g * 0.6, g, np.clip(g * 0.3 + 60, 0, 255)], -1).astype(np.uint8)


@pytest.mark.parametrize("tinted", [False, True])
@pytest.mark.parametrize("gs", [600, 4000])
def test_cfl_estimate_matches_the_oracle(ctx, tinted, gs):
    import jpxl_amd
    from jpxl_amd import _lib

    w, h = 320, 200
    rgb = _tinted(w, h, 23) if tinted else _rgb(w, h, 23)
    img = jpxl_amd.Image(ctx, rgb, 8)
    q = _quantizer(gs=gs)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL, True)
    img.quantize(q)
    plan = img.plan_download()
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    bmap = orc.cover(xyb, _qp(q))
    assert np.array_equal(plan.block_map, bmap)
    coeff = orc.forward_frame(xyb, bmap)
    cx, cb, lff = orc.estimate_cfl(w, h, bmap, coeff, _qp(q))
    assert np.array_equal(plan.cfl_x, cx) and np.array_equal(plan.cfl_b, cb)
    assert plan.lf_factors == lff
    if tinted and gs == 4000:
        assert np.abs(cb).max() > 0  # the tinted image correlates B with Y
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, _qp(q), cx, cb, lff)
    assert np.array_equal(plan.quant, oq) and np.array_equal(plan.lf, olf)


@pytest.mark.parametrize("wh", [(2100, 72), (96, 2060), (2056, 2056)])
def test_lists_across_lf_groups_match_the_oracle(ctx, wh):
    # Frames wider or taller than one 256x256-block LF group: the varblock
    # lists, CfL sample order and LF offsets run group after group, row after
    # row; structure, CfL and quantized planes against the oracle.
    import jpxl_amd
    from jpxl_amd import _lib

    w, h = wh
    rgb = _tinted(w, h, 41)
    img = jpxl_amd.Image(ctx, rgb, 8)
    q = _quantizer(gs=900)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL, True)
    img.quantize(q)
    plan = img.plan_download()
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    bmap = plan.block_map
    if w * h <= 300_000:
        assert np.array_equal(bmap, orc.cover(xyb, _qp(q)))
    coeff = orc.forward_frame(xyb, bmap)
    cx, cb, lff = orc.estimate_cfl(w, h, bmap, coeff, _qp(q))
    assert np.array_equal(plan.cfl_x, cx) and np.array_equal(plan.cfl_b, cb)
    assert plan.lf_factors == lff
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, _qp(q), cx, cb, lff)
    assert np.array_equal(plan.quant, oq) and np.array_equal(plan.lf, olf)


@pytest.mark.parametrize("epf,gab", [(0, 0), (1, 0), (2, 0), (3, 0), (1, 1), (3, 1)])
@pytest.mark.parametrize("mixed", [False, True])
def test_render_bit_exact(ctx, epf, gab, mixed):
    from jpxl_amd import Restoration, _lib

    w, h = 150, 77
    bmap = orc.mixed_map(w, h, seed=9) if mixed else orc.fixed_dct8_map(w, h)
    cfl = _cfl(bmap, 13) if mixed else None
    lff = (-3, 4) if mixed else (0, 0)
    img, xyb, lin = _image_and_oracle(ctx, w, h, 8, 8, bmap, cfl, lff)
    q = _quantizer(gs=1500)
    img.quantize(q)
    score, _, terms = img.render_score(Restoration(epf, gab, 7))
    got_xyb = img.download(_lib.BUF_RENDER_XYB)
    got_lin = img.download(_lib.BUF_LINEAR)
    plan = img.plan_download()
    rp = orc.RenderParams(q.global_scale, q.hf_mul, q.quant_lf, q.x_qm_scale, q.b_qm_scale, epf, gab, 7, 8)
    cx, cb = cfl if cfl else (None, None)
    olin, oxyb = orc.render(w, h, bmap, plan.quant, plan.lf, rp, cx, cb, lff)
    assert np.array_equal(got_xyb.view(np.uint32), oxyb.view(np.uint32))
    assert np.array_equal(got_lin.view(np.uint32), olin.view(np.uint32))
    oscore, oterms, _ = orc.ssimulacra2(lin, olin)
    assert abs(score - oscore) < SCORE_TOL
    _terms_close(terms, oterms)


@pytest.mark.parametrize("step", [0, 1, 2])
def test_epf_step_bit_exact(ctx, step):
    rng = np.random.default_rng(step)
    w, h = 67, 45
    planes = rng.uniform(-0.1, 0.6, (3, h, w)).astype(np.float32)
    sigma = rng.uniform(0.0, 3.0, ((h + 7) // 8, (w + 7) // 8)).astype(np.float32)
    got = ctx.epf_step(step, planes, sigma)
    want = orc.epf_step(step, planes, sigma)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_gaborish_bit_exact(ctx):
    rng = np.random.default_rng(1)
    planes = rng.uniform(0, 1, (3, 33, 70)).astype(np.float32)
    assert np.array_equal(ctx.gaborish(planes).view(np.uint32), orc.gaborish(planes).view(np.uint32))


@pytest.mark.parametrize("bits", [8, 16])
def test_metric_positive_xyb_matches_the_oracle(ctx, bits):
    # The metric's positive XYB of the reference image at scale 0 (its cube
    # root is certified against f32 rounding boundaries on the device) against
    # the oracle's correctly rounded root, sample by sample.
    import jpxl_amd
    from jpxl_amd import _lib

    w, h = 96, 64
    rgb = _rgb(w, h, 29, bits)
    img = jpxl_amd.Image(ctx, rgb, bits)
    got = img.download(_lib.BUF_REF_XYB0)
    _, lin, _ = orc.prepare_frame(rgb, bits)
    want = np.zeros_like(got)
    for y in range(h):
        for x in range(w):
            want[:, y, x] = orc.positive_xyb(lin[:, y, x])
    same = np.mean(got == want)
    ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert same >= 0.9999 and ulps.max() <= 1


def test_score_pair_identity_and_noise(ctx):
    rng = np.random.default_rng(2)
    ref = rng.uniform(0, 1, (3, 120, 100)).astype(np.float32)
    s, terms = ctx.score_pair(ref, ref)
    assert s == 100.0 and not terms.any()
    for amp in (0.02, 0.1):
        cand = np.clip(ref + rng.uniform(-amp, amp, ref.shape).astype(np.float32), 0, 1)
        s, terms = ctx.score_pair(ref, cand)
        os_, oterms, _ = orc.ssimulacra2(ref, cand)
        assert abs(s - os_) < PAIR_TOL
        _terms_close(terms, oterms, 1e-9)


@pytest.mark.parametrize("wh", [(8, 8), (9, 17), (256, 192)])
def test_score_pair_scale_counts(ctx, wh):
    w, h = wh
    ref = np.random.default_rng(3).uniform(0, 1, (3, h, w)).astype(np.float32)
    cand = np.clip(ref * 0.97 + 0.01, 0, 1)
    s, terms = ctx.score_pair(ref, cand)
    os_, oterms, _ = orc.ssimulacra2(ref, cand)
    assert len(terms) == len(oterms)
    assert abs(s - os_) < PAIR_TOL


@pytest.mark.parametrize("wh", [(1024, 768), (1531, 1027)])
def test_score_pair_long_lines_match_the_recursion(ctx, wh):
    # Long lines: the reference's recursion carries the rounding of every
    # f32 pair sum from the start of the line (an exact 9-tap blur drifts
    # from it by ~0.1 point at these sizes); the device carries the same.
    # Terms: an absolute bound for the rare outputs the reference's own f64
    # rounding sends to the other f32 neighbour.
    w, h = wh
    rng = np.random.default_rng(w)
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.3 * np.sin(xx / 37.0)[None] * np.cos(yy / 23.0)[None]
    ref = np.clip(base + rng.uniform(-0.05, 0.05, (3, h, w)), 0, 1).astype(np.float32)
    cand = np.clip(ref + rng.uniform(-0.02, 0.02, ref.shape).astype(np.float32), 0, 1)
    s, terms = ctx.score_pair(ref, cand)
    os_, oterms, _ = orc.ssimulacra2(ref, cand)
    assert abs(s - os_) < PAIR_TOL
    _terms_close(terms, oterms, 1e-7)


@pytest.mark.parametrize("case", [(1024, 768, 16, 5, 59199), (2048, 2048, 16, 2, 66525)])
def test_render_score_at_full_size(ctx, case):
    # The probe chain's score of a rendered candidate at the encode sizes
    # (2048 x 2048 RGB16 is the throughput workload) against the oracle's
    # SSIMULACRA2 of the same linear planes.
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    w, h, bits, seed, gs = case
    rgb = _rgb(w, h, seed, bits)
    img = jpxl_amd.Image(ctx, rgb, bits)
    bmap = orc.fixed_dct8_map(w, h)
    z = np.zeros(((bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8), np.int8)
    img.set_structure(bmap, z, z, (0, 0))
    img.quantize(Quantizer(gs, 1, 4, 3, 3, 1, 4.0, 1.0))
    score, _, terms = img.render_score(Restoration(1, 1, 7))
    got_lin = img.download(_lib.BUF_LINEAR)
    _, lin, _ = orc.prepare_frame(rgb, bits)
    oscore, oterms, _ = orc.ssimulacra2(lin, got_lin)
    assert abs(score - oscore) < SCORE_TOL
    _terms_close(terms, oterms)
    img.close()


def test_encode_quality_meets_target(ctx):
    import jpxl_amd

    rgb = _rgb(256, 192, 21)
    img = jpxl_amd.Image(ctx, rgb, 8)
    res = img.encode_quality(85.0, "balanced")
    # a first probe inside the accept band may end the search (trusted-first-probe)
    assert 1 <= res.probes <= 6
    assert res.met == 1 and res.score >= 85.0
    assert res.status in (0, 1, 2, 6)  # met, adjacent rungs, work cap, rescued
    # the chosen plan is left quantized: rescoring reproduces the score
    score, _, _ = img.render_score()
    assert score == res.score


def test_errors_are_typed(ctx):
    import jpxl_amd
    from jpxl_amd import JpxlgError

    with pytest.raises(JpxlgError):
        jpxl_amd.Image(ctx, np.zeros((4, 4, 3), np.uint8), 8)  # below the 8x8 floor
    img = jpxl_amd.Image(ctx, _rgb(16, 16, 1), 8)
    bad = np.full((2, 2), 16, np.uint8)  # overlapping varblocks
    with pytest.raises(JpxlgError):
        img.set_structure(bad)
    with pytest.raises(JpxlgError):
        img.quantize(_quantizer())  # no structure yet
    img.set_structure(orc.fixed_dct8_map(16, 16))
    with pytest.raises(JpxlgError):
        img.quantize(_quantizer(gs=0))


def test_full_size_config_properties(ctx):
    """2048x2048 RGB16: quantization bit-exact against the oracle, probe
    determinism, and a score inside the metric's range."""
    from jpxl_amd import Restoration

    w = h = 2048
    bmap = orc.fixed_dct8_map(w, h)
    img, xyb, lin = _image_and_oracle(ctx, w, h, 2, 16, bmap)
    q = _quantizer(gs=900)
    img.quantize(q)
    plan = img.plan_download()
    coeff = orc.forward_frame(xyb, bmap)
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, _qp(q))
    assert np.array_equal(plan.quant, oq) and np.array_equal(plan.lf, olf)
    s1 = img.render_score(Restoration(1, 0, 7))[0]
    img.quantize(q)
    s2 = img.render_score(Restoration(1, 0, 7))[0]
    # Global scale 900 is a very coarse quantizer, so the score may be
    # negative; it is finite, repeatable and below a finer plan's.
    assert s1 == s2 and np.isfinite(s1) and s1 < 100.0
    img.quantize(_quantizer(gs=20000))
    s3 = img.render_score(Restoration(1, 0, 7))[0]
    assert s1 < s3 < 100.0
    img.close()


def test_batch_probe_matches_one_image_at_a_time(ctx):
    """A batched probe (one launch per stage over uneven images of different
    depths, maps, CfL and quantizers) reproduces each image's own probe."""
    import jpxl_amd
    from jpxl_amd import Restoration

    specs = [((203, 141), 8, 7), ((64, 64), 16, 3), ((130, 97), 8, 11), ((300, 77), 12, 5)]
    imgs, qs = [], []
    for k, ((w, h), bits, seed) in enumerate(specs):
        bmap = orc.mixed_map(w, h, seed=seed)
        img, _, _ = _image_and_oracle(ctx, w, h, seed, bits, bmap, _cfl(bmap, seed), (k, -k))
        imgs.append(img)
        qs.append(_quantizer(gs=1500 + 700 * k, mul=1 + k % 2))
    r = Restoration(1, 0, 7)
    scores = jpxl_amd.batch_probe(imgs, qs, r)
    plans = [im.plan_download() for im in imgs]
    for k, im in enumerate(imgs):
        im.quantize(qs[k])
        s1 = im.render_score(r)[0]
        p1 = im.plan_download()
        assert np.array_equal(p1.quant, plans[k].quant) and np.array_equal(p1.lf, plans[k].lf)
        assert s1 == scores[k]
    for im in imgs:
        im.close()


def test_batch_encode_quality_matches_one_image_at_a_time(ctx):
    import jpxl_amd

    sizes = [(256, 192), (120, 200), (333, 111)]
    targets = [85.0, 70.0, 90.0]
    rgbs = [_rgb(w, h, 40 + k) for k, (w, h) in enumerate(sizes)]
    batch = [jpxl_amd.Image(ctx, x, 8) for x in rgbs]
    res = jpxl_amd.batch_encode_quality(batch, targets, "balanced")
    for k, x in enumerate(rgbs):
        one = jpxl_amd.Image(ctx, x, 8)
        r1 = one.encode_quality(targets[k], "balanced")
        assert (res[k].rung, res[k].probes, res[k].met) == (r1.rung, r1.probes, r1.met)
        assert res[k].score == r1.score
        assert [res[k].probe_rungs[i] for i in range(r1.probes)] == [r1.probe_rungs[i] for i in range(r1.probes)]
        assert np.array_equal(batch[k].plan_download().quant, one.plan_download().quant)
        one.close()
    for im in batch:
        im.close()


def test_server_results_are_the_same_alone_and_in_the_mix(ctx):
    """Continuous batching: requests submitted while others are mid-search
    join the running rounds; every result equals the request served alone."""
    import jpxl_amd

    sizes = [(256, 192), (150, 222), (300, 120), (96, 96), (200, 200)]
    targets = [85.0, 90.0, 70.0, 85.0, 90.0]
    efforts = ["balanced", "fast", "balanced", "fast", "balanced"]
    rgbs = [_rgb(w, h, 60 + k) for k, (w, h) in enumerate(sizes)]
    alone = []
    for k, x in enumerate(rgbs):
        im = jpxl_amd.Image(ctx, x, 8)
        r = im.encode_quality(targets[k], efforts[k])
        alone.append((r.rung, r.probes, r.score, im.plan_download().quant))
        im.close()
    srv = jpxl_amd.Server(ctx)
    imgs = [jpxl_amd.Image(ctx, x, 8) for x in rgbs]
    tickets = {}
    # two requests first, the rest arrive one per round
    for k in (0, 1):
        tickets[k] = srv.submit(imgs[k], targets[k], efforts[k])
    nxt, done = 2, {}
    while len(done) < len(rgbs):
        srv.step()
        for k, t in tickets.items():
            if k not in done:
                res = srv.poll(t)
                if res is not None:
                    done[k] = res
        if nxt < len(rgbs):
            tickets[nxt] = srv.submit(imgs[nxt], targets[nxt], efforts[nxt])
            nxt += 1
    for k in range(len(rgbs)):
        r = done[k]
        assert (r.rung, r.probes, r.score) == alone[k][:3]
        assert np.array_equal(imgs[k].plan_download().quant, alone[k][3])
        srv.release(tickets[k])
    assert srv.pending() == 0
    srv.close()
    for im in imgs:
        im.close()


def test_server_large_lane_round_outlives_many_small_rounds(ctx):
    """An image above 4 MP runs in the large-image lane; while its round is on
    the device, many small-lane rounds and hand-off quantizations take upload
    slots.  The large round's slot must stay its own until it is read back, so
    every result equals the request served alone."""
    import jpxl_amd

    big = (2080, 2064)  # 4.29 MP: the large lane
    smalls = [(96 + 8 * k, 80 + 4 * k) for k in range(12)]
    sizes = [big] + smalls
    targets = [85.0] + [70.0, 85.0, 90.0] * 4
    efforts = ["balanced"] + ["fast", "balanced"] * 6
    rgbs = [_rgb(w, h, 200 + k) for k, (w, h) in enumerate(sizes)]
    alone = []
    for k, x in enumerate(rgbs):
        im = jpxl_amd.Image(ctx, x, 8)
        r = im.encode_quality(targets[k], efforts[k])
        alone.append((r.rung, r.probes, r.score, im.plan_download().quant))
        im.close()
    srv = jpxl_amd.Server(ctx)
    imgs = [jpxl_amd.Image(ctx, x, 8) for x in rgbs]
    tickets = {0: srv.submit(imgs[0], targets[0], efforts[0])}
    nxt, done, steps = 1, {}, 0
    while len(done) < len(rgbs):
        srv.step()
        steps += 1
        for k, t in tickets.items():
            if k not in done:
                res = srv.poll(t)
                if res is not None:
                    done[k] = res
        # two small requests join per step so the small lane keeps cycling
        for _ in range(2):
            if nxt < len(rgbs):
                tickets[nxt] = srv.submit(imgs[nxt], targets[nxt], efforts[nxt])
                nxt += 1
    assert steps >= 9
    for k in range(len(rgbs)):
        r = done[k]
        assert (r.rung, r.probes, r.score) == alone[k][:3], k
        assert np.array_equal(imgs[k].plan_download().quant, alone[k][3]), k
        srv.release(tickets[k])
    srv.close()
    for im in imgs:
        im.close()


@pytest.mark.parametrize("wh", [(61, 37), (256, 130), (9, 11)])
def test_gaborish_precondition_bit_exact(ctx, wh):
    """prepare_gaborish_frame: the source XYB replaced by its 12-step Jacobi
    Gaborish inverse, bit for bit; the structure must be planned again."""
    import jpxl_amd
    from jpxl_amd import JpxlgError, _lib

    w, h = wh
    rgb = _rgb(w, h, 31)
    img = jpxl_amd.Image(ctx, rgb, 8)
    bmap = orc.fixed_dct8_map(w, h)
    img.set_structure(bmap)
    img.precondition_gaborish()
    got = img.download(_lib.BUF_XYB)
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    want = orc.gaborish_sharpen(xyb, 12)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    with pytest.raises(JpxlgError):
        img.quantize(_quantizer())
    img.set_structure(bmap)
    assert np.array_equal(img.download(_lib.BUF_COEFF).view(np.uint32),
                          orc.forward_frame(want, bmap).view(np.uint32))
    img.close()



AQ_MODES = [1, 2, 3, 5]  # masking, uniform, fine masking, edge refine


@pytest.mark.parametrize("mode", AQ_MODES)
def test_aq_field_matches_the_oracle(ctx, mode):
    import jpxl_amd
    from jpxl_amd import _lib

    w, h = 203, 141
    rgb = _rgb(w, h, 31)
    img = jpxl_amd.Image(ctx, rgb, 8)
    active = img.set_aq(mode)
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    adj, oactive, _ = orc.aq_field(xyb, mode)
    assert active == oactive
    if active:
        got = img.download(_lib.BUF_AQ_FIELD)
        assert np.array_equal(got.view(np.uint32), adj.view(np.uint32))


def test_aq_flat_frame_is_inactive(ctx):
    import jpxl_amd

    img = jpxl_amd.Image(ctx, np.full((64, 64, 3), 100, np.uint8), 8)
    assert not img.set_aq(1)


@pytest.mark.parametrize("mode", AQ_MODES)
@pytest.mark.parametrize("cover", [0, 1])
@pytest.mark.parametrize("gs,truncate", [(32768, 0), (3001, 1)])
def test_aq_plan_and_render_match_the_oracle(ctx, mode, cover, gs, truncate):
    """The fixed-quantizer plan with an AQ field (lib.rs:600-760): the wire
    quantizer, cover, CfL, per-varblock HfMul, quantized planes and the
    render with the per-varblock multiplier and EPF sigma."""
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    w, h = 203, 141
    rgb = _rgb(w, h, 31)
    img = jpxl_amd.Image(ctx, rgb, 8)
    active = img.set_aq(mode)
    assert active
    q = Quantizer(gs, 1, 16, 2, 2, truncate, 1.0, 1.0)
    img.plan_structure(q, cover, True)
    img.quantize(q)
    plan = img.plan_download()
    muls, wire, ro = img.plan_aq()
    xyb, lin, _ = orc.prepare_frame(rgb, 8)
    adj, _, fine = orc.aq_field(xyb, mode)
    gs2, qlf2, base, oro = orc.aq_setup(gs, 16, 1, fine)
    assert (wire.global_scale, wire.quant_lf, wire.hf_mul, ro) == (gs2, qlf2, base, oro)
    qp = orc.with_aq(orc.QuantParams(gs2, base, qlf2, 2, 2, truncate, 1.0, 1.0), adj, oro, fine)
    bmap = orc.cover(xyb, qp) if cover == 1 else orc.fixed_dct8_map(w, h)
    assert np.array_equal(plan.block_map, bmap)
    omuls = orc.aq_block_muls(w, h, bmap, qp)
    assert np.array_equal(muls, omuls)
    assert len(np.unique(omuls[bmap > 0])) > 1  # the field moves some varblocks off the baseline
    coeff = orc.forward_frame(xyb, bmap)
    cx, cb, lff = orc.estimate_cfl(w, h, bmap, coeff, qp)
    assert np.array_equal(plan.cfl_x, cx) and np.array_equal(plan.cfl_b, cb) and plan.lf_factors == lff
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, qp, cx, cb, lff)
    assert np.array_equal(plan.lf, olf)
    assert np.array_equal(plan.quant, oq)
    score, _, _ = img.render_score(Restoration(1, 0, 7))
    rp = orc.with_aq(orc.RenderParams(gs2, base, qlf2, 2, 2, 1, 0, 7, 8), adj, oro, fine)
    olin, oxyb = orc.render(w, h, bmap, oq, olf, rp, cx, cb, lff)
    assert np.array_equal(img.download(_lib.BUF_RENDER_XYB).view(np.uint32), oxyb.view(np.uint32))
    assert np.array_equal(img.download(_lib.BUF_LINEAR).view(np.uint32), olin.view(np.uint32))
    oscore, _, _ = orc.ssimulacra2(lin, olin)
    assert abs(score - oscore) < SCORE_TOL


def test_aq_refine_only_fallback(ctx):
    # an odd global_scale of 1 cannot be halved: AqSetup keeps the request's
    # quantizer and only refines (lib.rs:1899-1906)
    import jpxl_amd
    from jpxl_amd import Quantizer

    w, h = 96, 80
    rgb = _rgb(w, h, 33)
    img = jpxl_amd.Image(ctx, rgb, 8)
    assert img.set_aq(1)
    q = Quantizer(1, 1, 16, 2, 2, 0, 1.0, 1.0)
    img.plan_structure(q, 0, False)
    img.quantize(q)
    muls, wire, ro = img.plan_aq()
    assert ro and (wire.global_scale, wire.hf_mul, wire.quant_lf) == (1, 1, 16)
    assert muls[muls > 0].min() >= 1 and muls.max() <= 2
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    adj, _, _ = orc.aq_field(xyb, 1)
    qp = orc.with_aq(orc.QuantParams(1, 1, 16, 2, 2, 0, 1.0, 1.0), adj, True, False)
    bmap = orc.fixed_dct8_map(w, h)
    assert np.array_equal(muls, orc.aq_block_muls(w, h, bmap, qp))
    oq, olf = orc.quantize_frame(w, h, bmap, orc.forward_frame(xyb, bmap), qp)
    plan = img.plan_download()
    assert np.array_equal(plan.quant, oq) and np.array_equal(plan.lf, olf)


def test_graph_replays_match_the_eager_chain(ctx):
    """A captured chain replayed several times, with eager chains and
    synchronisations between the replays: every replay equals the eager
    batch probe of the same quantizers (scores, quantized planes, render)."""
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    sizes = [(203, 141), (96, 80), (130, 97)]
    imgs = []
    for i, (w, h) in enumerate(sizes):
        img = jpxl_amd.Image(ctx, _rgb(w, h, 40 + i), 8)
        img.plan_structure(_quantizer(gs=1500), _lib.COVER_HIERARCHICAL, True)
        imgs.append(img)
    r = Restoration(1, 0, 7)
    g = jpxl_amd.Graph(imgs, r)
    try:
        for rep, gs in enumerate([900, 2500, 900, 6000]):
            qs = [Quantizer(gs + 17 * k, 2, 4, 3, 3, 1, 4.0, 1.0) for k in range(len(imgs))]
            got = g.probe(qs)
            plans = [im.plan_download() for im in imgs]
            lins = [im.download(_lib.BUF_LINEAR) for im in imgs]
            # eager work between replays: the same probe through the batched chain
            want = jpxl_amd.batch_probe(imgs, qs, r)
            assert np.array_equal(got, want), (rep, got, want)
            for im, p, lin in zip(imgs, plans, lins):
                e = im.plan_download()
                assert np.array_equal(p.quant, e.quant) and np.array_equal(p.lf, e.lf)
                assert np.array_equal(lin.view(np.uint32), im.download(_lib.BUF_LINEAR).view(np.uint32))
            ctx.lib.jpxlg_profile_enable(ctx.handle, 0)  # a host call between replays
    finally:
        g.close()


def test_graph_rejects_a_changed_structure(ctx):
    import jpxl_amd
    from jpxl_amd import JpxlgError, Restoration, _lib

    img = jpxl_amd.Image(ctx, _rgb(96, 80, 44), 8)
    img.plan_structure(_quantizer(gs=1500), _lib.COVER_FIXED_DCT8, False)
    g = jpxl_amd.Graph([img], Restoration(1, 0, 7))
    try:
        g.probe([_quantizer(gs=1500)])
        img.plan_structure(_quantizer(gs=1500), _lib.COVER_HIERARCHICAL, False)
        with pytest.raises(JpxlgError):
            g.probe([_quantizer(gs=1500)])
    finally:
        g.close()


def test_autotune_keeps_the_scores(ctx):
    """Every launch shape the tuner may pick gives the same metric sums; the
    chosen one is reported and used by later chains of that size."""
    import jpxl_amd
    from jpxl_amd import Restoration, _lib

    w, h = 320, 200
    img = jpxl_amd.Image(ctx, _rgb(w, h, 50), 8)
    q = _quantizer(gs=1800)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL, True)
    img.quantize(q)
    before = img.render_score(Restoration(1, 0, 7))
    upw, sw = ctx.autotune(w, h)
    assert upw in (1, 2, 4) and sw in (0, 1)
    after = img.render_score(Restoration(1, 0, 7))
    assert before[0] == after[0] and np.array_equal(before[2], after[2])


@pytest.mark.parametrize("case", range(12))
def test_random_configurations_end_to_end(ctx, case):
    # Seeded sweep over frame sizes, bit depths, quantizers, cover modes, CfL,
    # AQ and restoration settings: structure, CfL, quantized planes and the
    # restored render bit-exact against the oracle, the score within the
    # metric tolerance.
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    rng = np.random.default_rng(1000 + case)
    w, h = int(rng.integers(9, 420)), int(rng.integers(9, 300))
    bits = (8, 8, 12, 16)[case % 4]
    gs = int(rng.choice([300, 900, 2000, 6000, 20000]))
    cover = int(rng.choice([_lib.COVER_FIXED_DCT8, _lib.COVER_HIERARCHICAL]))
    use_cfl = case % 3 != 0
    truncate = int(rng.integers(0, 2))
    aq = (orc.AQ_MASKING, orc.AQ_FINE_MASKING, orc.AQ_EDGE_REFINE)[case % 3] if bits == 8 else orc.AQ_OFF
    epf, gab = int(rng.integers(0, 4)), int(rng.integers(0, 2))
    print(f"case {case}: {w}x{h} bits {bits} gs {gs} cover {cover} cfl {use_cfl} trunc {truncate} aq {aq} "
          f"epf {epf} gab {gab}", flush=True)
    rgb = _rgb(w, h, 500 + case, bits)
    img = jpxl_amd.Image(ctx, rgb, bits)
    xyb, lin, _ = orc.prepare_frame(rgb, bits)
    q = Quantizer(gs, 1, 4, 3, 3, truncate, 4.0, 1.0)
    qp = _qp(q)
    rp = orc.RenderParams(gs, 1, 4, 3, 3, epf, gab, 7, bits)
    if aq != orc.AQ_OFF and img.set_aq(aq):
        q = Quantizer(gs, 1, 16, 2, 2, truncate, 1.0, 1.0)
        img.plan_structure(q, cover, use_cfl)
        img.quantize(q)
        _, wire, ro = img.plan_aq()
        adj, _, fine = orc.aq_field(xyb, aq)
        gs2, qlf2, base, oro = orc.aq_setup(gs, 16, 1, fine)
        assert (wire.global_scale, wire.quant_lf, wire.hf_mul, ro) == (gs2, qlf2, base, oro)
        qp = orc.with_aq(orc.QuantParams(gs2, base, qlf2, 2, 2, truncate, 1.0, 1.0), adj, oro, fine)
        rp = orc.with_aq(orc.RenderParams(gs2, base, qlf2, 2, 2, epf, gab, 7, bits), adj, oro, fine)
    else:
        img.plan_structure(q, cover, use_cfl)
    img.quantize(q)
    plan = img.plan_download()
    bmap = orc.cover(xyb, qp) if cover == _lib.COVER_HIERARCHICAL else orc.fixed_dct8_map(w, h)
    assert np.array_equal(plan.block_map, bmap)
    coeff = orc.forward_frame(xyb, bmap)
    if use_cfl:
        cx, cb, lff = orc.estimate_cfl(w, h, bmap, coeff, qp)
    else:
        cx = cb = np.zeros_like(plan.cfl_x)
        lff = (0, 0)
    assert np.array_equal(plan.cfl_x, cx) and np.array_equal(plan.cfl_b, cb) and plan.lf_factors == lff
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, qp, cx, cb, lff)
    assert np.array_equal(plan.lf, olf) and np.array_equal(plan.quant, oq)
    score, _, _ = img.render_score(Restoration(epf, gab, 7))
    olin, oxyb = orc.render(w, h, bmap, oq, olf, rp, cx, cb, lff)
    assert np.array_equal(img.download(_lib.BUF_RENDER_XYB).view(np.uint32), oxyb.view(np.uint32))
    assert np.array_equal(img.download(_lib.BUF_LINEAR).view(np.uint32), olin.view(np.uint32))
    oscore, _, _ = orc.ssimulacra2(lin, olin)
    assert abs(score - oscore) < SCORE_TOL
    img.close()


@pytest.mark.parametrize("aq", [False, True])
@pytest.mark.parametrize("epf", [0, 1, 3])
def test_fused_quantize_render_matches_the_separate_stages(ctx, aq, epf):
    # A probe chain quantizes and reconstructs the exact DCT form in one
    # launch (the reconstruction works from the levels in LDS); its planes,
    # levels and score must equal the separate quantize and render stages,
    # and the render must equal the oracle's.
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    w, h = 237, 165
    rgb = _rgb(w, h, 77)
    q = Quantizer(1200, 1, 16 if aq else 4, 2 if aq else 3, 2 if aq else 3, 1, 1.0 if aq else 4.0, 1.0)
    r = Restoration(epf, 0, 7)
    imgs = []
    for _ in range(2):
        img = jpxl_amd.Image(ctx, rgb, 8)
        if aq:
            assert img.set_aq(orc.AQ_MASKING)
        img.plan_structure(q, _lib.COVER_HIERARCHICAL, True)
        imgs.append(img)
    fused_score = jpxl_amd.batch_probe([imgs[0]], [q], r)[0]
    imgs[1].quantize(q)
    sep_score, _, _ = imgs[1].render_score(r)
    assert fused_score == sep_score
    for buf in (_lib.BUF_RENDER_XYB, _lib.BUF_LINEAR, _lib.BUF_SIGMA):
        assert np.array_equal(imgs[0].download(buf).view(np.uint32), imgs[1].download(buf).view(np.uint32))
    p0, p1 = imgs[0].plan_download(), imgs[1].plan_download()
    assert np.array_equal(p0.quant, p1.quant) and np.array_equal(p0.lf, p1.lf)
    if not aq:
        rp = orc.RenderParams(q.global_scale, q.hf_mul, q.quant_lf, q.x_qm_scale, q.b_qm_scale, epf, 0, 7, 8)
        olin, oxyb = orc.render(w, h, p1.block_map, p1.quant, p1.lf, rp, p1.cfl_x, p1.cfl_b, p1.lf_factors)
        assert np.array_equal(imgs[0].download(_lib.BUF_RENDER_XYB).view(np.uint32), oxyb.view(np.uint32))
    for img in imgs:
        img.close()
