"""The drop-in boundary on the GPU, through the C ABI: a plan built on the
host is rendered and scored exactly (PerceptualEvaluator::evaluate over a
ValidatedPixelPlan, jpxl-encode-policy/src/quality.rs:87-157), and the
quantized plan is handed over in the QuantizedFrameIr order
(jpxl-encode/src/vardct/plan.rs:519-636)."""
import numpy as np
import pytest

import oracle_binding as orc

pytestmark = pytest.mark.gpu

SCORE_TOL = 2e-4


@pytest.fixture(scope="module")
def ctx():
    import jpxl_amd

    c = jpxl_amd.Context(0)
    yield c
    c.close()


def _rgb(w, h, seed, bits=8):
    from jpxl_amd.synth import synthetic_rgb

    return synthetic_rgb(w, h, seed, bits)


def _cfl(bmap, seed):
    th, tw = (bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8
    rng = np.random.default_rng(seed)
    return rng.integers(-20, 21, (th, tw)).astype(np.int8), rng.integers(-20, 21, (th, tw)).astype(np.int8)


def ir_from_footprint(quant, lf, bmap):
    """QuantizedFrameIr restated from its definition (plan.rs:434-636) over
    the footprint layout: LF groups of 256 x 256 blocks in raster order; in a
    group, varblocks by the block-raster order of their top-left blocks, each
    with its X, Y, B arrays in the coefficient array's raster order (cell k of
    a side-s square at row 8 by + k / s, column 8 bx + k % s); each group's LF
    planes in raster order."""
    bh, bw = bmap.shape
    arena, lfg, vbs = [], [], []
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
                    start = sum(len(a) for a in arena)
                    for c in range(3):
                        arena.append(quant[c, 8 * by:8 * by + s, 8 * bx:8 * bx + s].reshape(-1))
                    vbs.append(dict(lf_group=gi, bx=bx, by=by, side=s, start=start))
    return np.concatenate(arena), np.concatenate(lfg), vbs


@pytest.mark.parametrize("wh", [(203, 141), (2100, 2070)])
def test_ir_hand_off_equals_the_footprint_regathered(ctx, wh):
    import jpxl_amd
    from jpxl_amd import Quantizer

    w, h = wh
    rgb = _rgb(w, h, 51)
    img = jpxl_amd.Image(ctx, rgb, 8)
    bmap = orc.mixed_map(w, h, seed=5)
    cx, cb = _cfl(bmap, 3)
    img.set_structure(bmap, cx, cb, (3, -4))
    q = Quantizer(3000, 2, 4, 3, 3, 1, 4.0, 1.0)
    img.quantize(q)
    plan = img.plan_download()
    arena, lfg, vbs = img.plan_download_ir()
    want_arena, want_lfg, want_vbs = ir_from_footprint(plan.quant, plan.lf, plan.block_map)
    assert np.array_equal(arena, want_arena)
    assert np.array_equal(lfg, want_lfg)
    assert [(v["lf_group"], v["bx"], v["by"], v["side"], v["start"]) for v in vbs] == \
        [(v["lf_group"], v["bx"], v["by"], v["side"], v["start"]) for v in want_vbs]
    assert all(v["hf_mul"] == 2 for v in vbs)
    if w > 2048:
        assert len({v["lf_group"] for v in vbs}) == 4
    img.close()


@pytest.mark.parametrize("mixed,epf", [(False, 1), (True, 1), (True, 3)])
def test_a_host_plan_renders_bit_exact(ctx, mixed, epf):
    """A plan quantized by the oracle on the host, uploaded as a whole, renders
    the oracle's pixels bit for bit and scores within the metric tolerance;
    the device hands the same plan back."""
    import jpxl_amd
    from jpxl_amd import Plan, Quantizer, Restoration, _lib

    w, h = 203, 141
    rgb = _rgb(w, h, 52)
    xyb, lin, _ = orc.prepare_frame(rgb, 8)
    bmap = orc.mixed_map(w, h, seed=9) if mixed else orc.fixed_dct8_map(w, h)
    cx, cb = _cfl(bmap, 4) if mixed else (np.zeros(((bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8), np.int8),) * 2
    lff = (5, -9) if mixed else (0, 0)
    q = Quantizer(2400, 1, 4, 3, 3, 1, 4.0, 1.0)
    qp = orc.QuantParams(q.global_scale, q.hf_mul, q.quant_lf, 3, 3, 1, 4.0, 1.0)
    coeff = orc.forward_frame(xyb, bmap)
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, qp, cx, cb, lff)
    img = jpxl_amd.Image(ctx, rgb, 8)
    img.set_plan(q, Plan(oq, olf, bmap, cx, cb, lff))
    back = img.plan_download()
    assert np.array_equal(back.quant, oq) and np.array_equal(back.lf, olf) and np.array_equal(back.block_map, bmap)
    score, _, _ = img.render_score(Restoration(epf, 0, 7))
    rp = orc.RenderParams(q.global_scale, q.hf_mul, q.quant_lf, 3, 3, epf, 0, 7, 8)
    olin, oxyb = orc.render(w, h, bmap, oq, olf, rp, cx, cb, lff)
    assert np.array_equal(img.download(_lib.BUF_LINEAR).view(np.uint32), olin.view(np.uint32))
    oscore, _, _ = orc.ssimulacra2(lin, olin)
    assert abs(score - oscore) < SCORE_TOL
    img.close()


def test_a_host_plan_with_per_varblock_hf_mul(ctx):
    """Per-varblock HfMul supplied with the plan (here the fixed-quantizer AQ
    field's, lib.rs:736-750) renders as the oracle renders that field; the IR
    table carries the same multipliers; quantizing again drops the plan's."""
    import jpxl_amd
    from jpxl_amd import Plan, Quantizer, Restoration, _lib

    w, h = 203, 141
    rgb = _rgb(w, h, 31)
    xyb, lin, _ = orc.prepare_frame(rgb, 8)
    adj, active, fine = orc.aq_field(xyb, _lib.AQ_MASKING)
    assert active
    gs2, qlf2, base, oro = orc.aq_setup(2500, 16, 1, fine)
    qp = orc.with_aq(orc.QuantParams(gs2, base, qlf2, 2, 2, 0, 1.0, 1.0), adj, oro, fine)
    bmap = orc.mixed_map(w, h, seed=12)
    muls = orc.aq_block_muls(w, h, bmap, qp)
    assert len(np.unique(muls[bmap > 0])) > 1
    coeff = orc.forward_frame(xyb, bmap)
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, qp)
    z = np.zeros(((bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8), np.int8)
    img = jpxl_amd.Image(ctx, rgb, 8)
    wire = Quantizer(gs2, base, qlf2, 2, 2, 0, 1.0, 1.0)
    img.set_plan(wire, Plan(oq, olf, bmap, z, z, (0, 0)), muls)
    got_muls, got_wire, ro = img.plan_aq()
    assert np.array_equal(got_muls, muls) and got_wire.global_scale == gs2 and not ro
    _, _, vbs = img.plan_download_ir()
    assert all(v["hf_mul"] == muls[v["by"], v["bx"]] for v in vbs)
    score, _, _ = img.render_score(Restoration(1, 0, 7))
    rp = orc.with_aq(orc.RenderParams(gs2, base, qlf2, 2, 2, 1, 0, 7, 8), adj, oro, fine)
    olin, _ = orc.render(w, h, bmap, oq, olf, rp)
    assert np.array_equal(img.download(_lib.BUF_LINEAR).view(np.uint32), olin.view(np.uint32))
    oscore, _, _ = orc.ssimulacra2(lin, olin)
    assert abs(score - oscore) < SCORE_TOL
    # quantizing replaces the supplied plan: its multipliers are gone
    img.quantize(wire)
    again, _, _ = img.plan_aq()
    assert set(np.unique(again[bmap > 0])) == {base}
    img.close()


def test_a_bad_host_plan_is_refused(ctx):
    import jpxl_amd
    from jpxl_amd import JpxlgError, Plan, Quantizer

    w, h = 64, 48
    img = jpxl_amd.Image(ctx, _rgb(w, h, 1), 8)
    bmap = orc.fixed_dct8_map(w, h)
    z = np.zeros((1, 1), np.int8)
    q = np.zeros((3, 48, 64), np.int32)
    lf = np.zeros((3, 6, 8), np.int32)
    q[1, 5, 5] = (1 << 20) + 1
    with pytest.raises(JpxlgError):
        img.set_plan(Quantizer(2000, 1, 4, 2, 2, 0, 1.0, 1.0), Plan(q, lf, bmap, z, z, (0, 0)))
    q[1, 5, 5] = 0
    muls = np.zeros((6, 8), np.uint32)
    with pytest.raises(JpxlgError):
        img.set_plan(Quantizer(2000, 1, 4, 2, 2, 0, 1.0, 1.0), Plan(q, lf, bmap, z, z, (0, 0)), muls)
    img.close()
