"""The codestream of GPU plans, through the C ABI: jpxlg_encode_emit writes
the image's last quantized plan (attach_entropy + emit_codestream_with_executor,
jpxl-encode-policy/src/lib.rs:1012, jpxl-encode/src/vardct/write.rs:350), and
the checker (a restatement of the reference decoder's reading rules,
tests/host/codestream_check.cpp) must read back exactly that plan: the
quantized HF coefficients, the LF planes, the block map, each varblock's
HfMul, the CfL tiles and LF factors, and the frame fields (wire quantizer,
restoration, bit depth).  On the quality path the plan read back is also
rendered by the oracle and scored: the stream decodes to the pixels the
search scored."""
import ctypes

import numpy as np
import pytest

import codestream_binding as cs
import oracle_binding as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import jpxl_amd

    c = jpxl_amd.Context(0)
    yield c
    c.close()


def _rgb(w, h, seed, bits=8):
    from jpxl_amd.synth import synthetic_rgb

    return synthetic_rgb(w, h, seed, bits)


def _assert_plan(d, img, restoration=None):
    plan = img.plan_download()
    muls, wire, _ = img.plan_aq()
    assert d.info.stage == 7
    assert (d.info.width, d.info.height) == (img.width, img.height)
    assert np.array_equal(d.block_map, plan.block_map)
    assert np.array_equal(d.quant, cs.llf_zeroed(plan.quant, plan.block_map))
    assert np.array_equal(d.lf, plan.lf)
    assert np.array_equal(d.hf_mul, np.where(plan.block_map > 0, muls, 0))
    assert np.array_equal(d.cfl_x, plan.cfl_x) and np.array_equal(d.cfl_b, plan.cfl_b)
    assert (d.info.x_factor_lf, d.info.b_factor_lf) == plan.lf_factors
    assert (d.info.global_scale, d.info.quant_lf, d.info.x_qm_scale, d.info.b_qm_scale) == (
        wire.global_scale, wire.quant_lf, wire.x_qm_scale, wire.b_qm_scale)
    if restoration is not None:
        assert (d.info.epf_iters, d.info.gaborish) == (restoration.epf_iters, restoration.gaborish)
        assert (d.sharpness == restoration.sharpness).all()
    return plan, wire


def test_quality_encode_stream_decodes_to_the_scored_pixels(ctx):
    """512x512 RGB8 at quality 85: the stream carries the chosen plan, and the
    plan read back, rendered by the oracle, scores what the search scored."""
    import jpxl_amd
    from jpxl_amd import _lib

    w = h = 512
    rgb = _rgb(w, h, 1)
    img = jpxl_amd.Image(ctx, rgb, 8)
    res = img.encode_quality(85.0, "balanced")
    stream, sizes = img.encode_emit()
    assert stream[:2] == b"\xff\x0a" and sizes.total == len(stream)
    d = cs.decode(stream)
    plan, wire = _assert_plan(d, img)
    assert (d.info.global_scale, d.info.quant_lf) == (res.global_scale, res.quant_lf)
    # render the plan the stream carries with the oracle and score it
    xyb, lin, _ = orc.prepare_frame(rgb, 8)
    rp = orc.RenderParams(d.info.global_scale, res.hf_mul, d.info.quant_lf, d.info.x_qm_scale,
                          d.info.b_qm_scale, d.info.epf_iters, d.info.gaborish, int(d.sharpness[0, 0]), 8)
    q = d.quant.copy()
    bm = d.block_map
    for by, bx in zip(*np.nonzero(bm)):  # LLF cells are carried by the LF planes; take the plan's
        c = int(bm[by, bx]) // 8
        q[:, 8 * by:8 * by + c, 8 * bx:8 * bx + c] = plan.quant[:, 8 * by:8 * by + c, 8 * bx:8 * bx + c]
    olin, _ = orc.render(w, h, bm, q, d.lf, rp, d.cfl_x, d.cfl_b, (d.info.x_factor_lf, d.info.b_factor_lf))
    oscore, _, _ = orc.ssimulacra2(lin, olin)
    assert abs(oscore - res.score) < 2e-4, (oscore, res.score)
    assert oscore >= 85.0 or not res.met
    img.close()


def test_fixed_quantizer_stream_carries_per_varblock_muls(ctx):
    """The fixed-quantizer request (AQ masking: per-varblock HfMul on the wire
    quantizer) over the hierarchical cover, at 12 bits."""
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    w, h = 640, 360
    img = jpxl_amd.Image(ctx, _rgb(w, h, 4, 12), 12)
    assert img.set_aq(_lib.AQ_MASKING)
    q = Quantizer(20000, 1, 16, 2, 2, 0, 1.0, 1.0)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL_LEGACY, True)
    img.quantize(q)
    r = Restoration(2, 1, 5)
    stream, _ = img.encode_emit(r)
    d = cs.decode(stream)
    _assert_plan(d, img, r)
    assert d.info.bits == 12
    assert len(np.unique(d.hf_mul[d.block_map > 0])) > 1
    img.close()


def test_frames_over_several_lf_groups(ctx):
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    w, h = 2100, 300
    img = jpxl_amd.Image(ctx, _rgb(w, h, 6), 8)
    q = Quantizer(30000, 1, 16, 3, 2, 1, 4.0, 0.0)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL, True)
    img.quantize(q)
    r = Restoration(1, 0, 7)
    stream, sizes = img.encode_emit(r, threads=4)
    d = cs.decode(stream)
    assert d.info.num_lf_groups == 2
    _assert_plan(d, img, r)
    img.close()


def test_full_size_quality_encode_stream(ctx):
    """The 2048x2048 RGB16 quality-85 encode: the chosen plan's stream reads
    back exactly; its size is the encode's byte count."""
    import jpxl_amd

    w = h = 2048
    img = jpxl_amd.Image(ctx, _rgb(w, h, 2, 16), 16)
    res = img.encode_quality(85.0, "balanced")
    stream, sizes = img.encode_emit()
    d = cs.decode(stream)
    _assert_plan(d, img)
    assert d.info.bits == 16
    assert sizes.total == len(stream)
    assert 0.2 < 8 * len(stream) / (w * h) < 4.0  # bits per pixel of a q85 photograph-like image
    img.close()


def test_facade_returns_a_codestream(ctx):
    """Encoder::with_quality(..).encode_rgb8 returns the codestream
    (jpxl/src/lib.rs:507-794); the report's byte count is its length."""
    import jpxl_amd

    w, h = 320, 240
    rgb = _rgb(w, h, 3)
    enc = jpxl_amd.Encoder(ctx).with_quality(80).with_effort("fast").with_threads(2)
    stream, report = enc.encode_rgb8_reported(w, h, rgb.tobytes())
    assert isinstance(stream, bytes) and stream[:2] == b"\xff\x0a"
    assert report.bytes == len(stream) and report.met
    d = cs.decode(stream)
    assert (d.info.width, d.info.height, d.info.global_scale) == (w, h, report.global_scale)
    fixed = jpxl_amd.Encoder(ctx).with_global_scale(32768).encode_rgb8(w, h, rgb)
    assert cs.decode(fixed).info.stage == 7


def test_priced_search_chooses_the_smallest_stream(ctx):
    """Mode 2 (the reference's default build, quality.rs:1886-2010): with two
    feasible finalists both are priced by the codestream stage and the chosen
    plan's stream has the priced size; with one, the choice is mode 0's."""
    import jpxl_amd
    from jpxl_amd import _lib

    priced = 0
    for k, (w, h, t) in enumerate([(320, 240, 85.0), (256, 256, 90.0), (400, 300, 70.0), (512, 384, 80.0),
                                   (300, 200, 88.0), (384, 384, 75.0)]):
        rgb = _rgb(w, h, 20 + k)
        img = jpxl_amd.Image(ctx, rgb, 8)
        res = img.encode_quality(t, "balanced", mode=_lib.MODE_PRICED)
        stream, sizes = img.encode_emit()
        assert cs.decode(stream).info.stage == 7
        plain = jpxl_amd.Image(ctx, rgb, 8)
        r0 = plain.encode_quality(t, "balanced")
        if res.exact_prices:
            priced += 1
            assert res.exact_prices == 2 and res.bytes == len(stream)
            s0, _ = plain.encode_emit()
            assert len(stream) <= len(s0)  # never a larger stream than the coarsest finalist's
        else:
            assert res.bytes == 0 and res.rung == r0.rung
        plain.close()
        img.close()


def test_byte_target_search_lands_under_the_target(ctx):
    """The byte-rate search (rate.rs:2016-2140 over search_ladder): every
    candidate priced exactly; the chosen plan's stream is the largest priced
    size at or under the target and is what jpxlg_encode_emit writes."""
    import jpxl_amd

    w, h = 480, 320
    rgb = _rgb(w, h, 11)
    img = jpxl_amd.Image(ctx, rgb, 8)
    for target in (6000, 20000, 45000):
        res = img.encode_bytes(target)
        stream, _ = img.encode_emit()
        trace = [(res.trace_rungs[i], res.trace_bytes[i]) for i in range(min(res.prices, 64))]
        assert 0 < len(stream) <= target and len(stream) == res.bytes
        assert res.bytes == max(b for _, b in trace if b <= target)
        assert res.prices <= 24
        d = cs.decode(stream)
        assert (d.info.global_scale, d.info.quant_lf) == (res.global_scale, res.quant_lf)
    img.close()


def test_byte_target_below_the_floor_is_refused(ctx):
    import jpxl_amd
    from jpxl_amd import JpxlgError

    img = jpxl_amd.Image(ctx, _rgb(200, 150, 12), 8)
    with pytest.raises(JpxlgError):
        img.encode_bytes(40)
    img.close()


def test_facade_target_bytes_and_bpp(ctx):
    import jpxl_amd

    w, h = 320, 240
    rgb = _rgb(w, h, 13)
    s1 = jpxl_amd.Encoder(ctx).with_target_bytes(12000).encode_rgb8(w, h, rgb)
    assert len(s1) <= 12000 and cs.decode(s1).info.stage == 7
    s2 = jpxl_amd.Encoder(ctx).with_target_bpp(1.0).encode_rgb8(w, h, rgb)
    assert len(s2) <= w * h // 8 and cs.decode(s2).info.stage == 7


@pytest.mark.parametrize("aq,r", [(True, (2, 1, 5)), (False, (3, 1, 2)), (True, (0, 0, 7))])
def test_device_render_is_the_decoded_stream(ctx, aq, r):
    """The device's rendered XYB planes (jpxlg_render_score, read back as
    JPXLG_BUF_RENDER_XYB) equal the planes the fixture-graded reconstruction
    (tests/vardct_render.py, within 4.6e-5 of the reference's djxl decodes)
    draws from the stream jpxlg_encode_emit writes for the same plan:
    per-varblock HfMul, CfL, Gaborish and EPF included."""
    import jpxl_amd
    import vardct_render as vr
    from jpxl_amd import Quantizer, Restoration, _lib

    w, h = 480, 272
    img = jpxl_amd.Image(ctx, _rgb(w, h, 31), 8)
    if aq:
        assert img.set_aq(_lib.AQ_MASKING)
    q = Quantizer(12000, 2, 16, 3, 2, 1, 1.0, 0.0)
    img.plan_structure(q, _lib.COVER_HIERARCHICAL, True)
    img.quantize(q)
    rest = Restoration(*r)
    img.render_score(rest)
    device = img.download(_lib.BUF_RENDER_XYB)
    stream, _ = img.encode_emit(rest)
    full = cs.decode_full(stream)
    if aq:
        assert len({v[3] for v in full.varblocks}) > 1
    ours = vr.render_xyb(full)
    err = np.abs(ours - device).max()
    assert err < 5e-6, err
    img.close()


@pytest.mark.parametrize("w,h,bits", [(512, 512, 8), (2100, 300, 8), (2048, 2048, 16)])
def test_device_runs_are_the_runs_of_the_ir(ctx, w, h, bits):
    """jpxlg_plan_download_runs: the device's runs equal those restated from
    jpxlg_plan_download_ir's arena, and both hand-offs give one stream."""
    import jpxl_amd

    img = jpxl_amd.Image(ctx, _rgb(w, h, 40, bits), bits)
    img.encode_quality(85.0, "balanced")
    arena, lfg, vbs = img.plan_download_ir()
    ra, runs, lfg2, vbs2 = img.plan_download_runs()
    assert vbs2 == vbs and np.array_equal(lfg2, lfg)
    want_arena, want = cs.runs_from_ir(arena, vbs)
    assert [(r.start, r.count, r.nonzeros) for r in runs] == want
    assert np.array_equal(ra, want_arena)
    from jpxl_amd import _lib

    frame = _lib.FrameDesc()
    ctx.check(ctx.lib.jpxlg_plan_frame_desc(img.handle, None, ctypes.byref(frame)))
    plan = img.plan_download()
    a, _ = cs.emit(frame, arena, lfg, vbs, plan.cfl_x, plan.cfl_b)
    b, _ = cs.emit_runs(frame, ra, runs, lfg, vbs, plan.cfl_x, plan.cfl_b)
    assert a == b
    c, _ = img.encode_emit()  # the library's own path goes through the runs
    assert c == a
    img.close()
