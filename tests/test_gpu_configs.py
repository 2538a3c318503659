"""The benchmark configurations at their own sizes, on the GPU, through the C
ABI: the plain 512x512 fixed-scale DCT8 encode against the oracle, the
2048x2048 cover against the oracle, the 4096x4096 precision forms against
the float64 DCT, 256 uneven images batched against one at a time, and a mix
of served requests against each request served alone.  The images come from
the configurations' generator (jpxl_amd.synth, the 1/f^2 field with shapes
and noise)."""
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


def test_plain_512_fixed_scale_dct8_matches_the_oracle(ctx):
    """One 512x512 RGB8 image (seed 1) at a fixed global scale with the DCT8
    AC strategy: the fixed-quantizer request's AQ masking field, CfL,
    quantized planes, render and SSIMULACRA2, each against the oracle."""
    import jpxl_amd
    from jpxl_amd import Quantizer, Restoration, _lib

    w = h = 512
    rgb = _rgb(w, h, 1)
    img = jpxl_amd.Image(ctx, rgb, 8)
    assert img.set_aq(_lib.AQ_MASKING)
    q = Quantizer(32768, 1, 16, 2, 2, 0, 1.0, 1.0)
    img.plan_structure(q, _lib.COVER_FIXED_DCT8, True)
    img.quantize(q)
    plan = img.plan_download()
    muls, wire, ro = img.plan_aq()
    xyb, lin, _ = orc.prepare_frame(rgb, 8)
    adj, _, fine = orc.aq_field(xyb, _lib.AQ_MASKING)
    gs2, qlf2, base, oro = orc.aq_setup(32768, 16, 1, fine)
    assert (wire.global_scale, wire.quant_lf, wire.hf_mul, ro) == (gs2, qlf2, base, oro)
    qp = orc.with_aq(orc.QuantParams(gs2, base, qlf2, 2, 2, 0, 1.0, 1.0), adj, oro, fine)
    bmap = orc.fixed_dct8_map(w, h)
    assert np.array_equal(plan.block_map, bmap)
    assert np.array_equal(muls, orc.aq_block_muls(w, h, bmap, qp))
    coeff = orc.forward_frame(xyb, bmap)
    cx, cb, lff = orc.estimate_cfl(w, h, bmap, coeff, qp)
    assert np.array_equal(plan.cfl_x, cx) and np.array_equal(plan.cfl_b, cb) and plan.lf_factors == lff
    oq, olf = orc.quantize_frame(w, h, bmap, coeff, qp, cx, cb, lff)
    assert np.array_equal(plan.quant, oq) and np.array_equal(plan.lf, olf)
    score, _, terms = img.render_score(Restoration(1, 0, 7))
    rp = orc.with_aq(orc.RenderParams(gs2, base, qlf2, 2, 2, 1, 0, 7, 8), adj, oro, fine)
    olin, _ = orc.render(w, h, bmap, oq, olf, rp, cx, cb, lff)
    assert np.array_equal(img.download(_lib.BUF_LINEAR).view(np.uint32), olin.view(np.uint32))
    oscore, oterms, _ = orc.ssimulacra2(lin, olin)
    assert abs(score - oscore) < SCORE_TOL
    assert np.abs(terms - oterms).max() < 1e-6
    img.close()


def test_cover_matches_the_oracle_at_2048(ctx):
    """The hierarchical cover of the 2048x2048 RGB16 image (seed 2) at a
    quality-search quantizer: the block map equals the oracle's."""
    import jpxl_amd
    from jpxl_amd import _lib

    w = h = 2048
    rgb = _rgb(w, h, 2, 16)
    img = jpxl_amd.Image(ctx, rgb, 16)
    q = jpxl_amd.quality_quantizer((3000, 1), "balanced")
    img.plan_structure(q, _lib.COVER_HIERARCHICAL, False)
    got = img.plan_download_map()
    xyb, _, _ = orc.prepare_frame(rgb, 16)
    qp = orc.QuantParams(q.global_scale, q.hf_mul, q.quant_lf, q.x_qm_scale, q.b_qm_scale, q.truncate,
                         q.lambda_scale, q.zero_token_bits)
    want = orc.cover(xyb, qp)
    assert np.array_equal(got, want)
    assert set(np.unique(want)) == {0, 8, 16, 32}
    img.close()


def test_precision_forms_at_4096(ctx):
    """4096x4096 RGB8 (seed 3) on its q85 plan: the FP32 and bf16 hi/lo
    matrix-core DCTs against the float64 closed form per transform size,
    their quantized planes against the exact path's, and the SSIMULACRA2
    delta at the same quantizer."""
    import jpxl_amd
    from jpxl_amd import _lib
    from jpxl_amd.precision import error_report, forward_frame_f64

    w = h = 4096
    img = jpxl_amd.Image(ctx, _rgb(w, h, 3), 8)
    res = img.encode_quality(85.0, "balanced")
    q = jpxl_amd.quality_quantizer((int(res.global_scale), int(res.hf_mul)), "balanced", int(res.quant_lf))
    plan0 = img.plan_download()
    bmap = plan0.block_map
    assert set(np.unique(bmap)) == {0, 8, 16, 32}
    want64 = forward_frame_f64(img.download(_lib.BUF_XYB), bmap)
    for form, tol in ((_lib.DCT_FP32, 2e-6), (_lib.DCT_BF16X3, 1e-4)):
        img.set_dct_form(form)
        rep = error_report(img.download(_lib.BUF_COEFF), want64, bmap)
        assert set(rep) == {"8", "16", "32"} and all(r["max_abs"] < tol for r in rep.values()), rep
        img.quantize(q)
        plan = img.plan_download()
        differ = int((plan.quant != plan0.quant).sum()) + int((plan.lf != plan0.lf).sum())
        assert differ <= 1e-4 * (plan.quant.size + plan.lf.size)
        score, _, _ = img.render_score()
        assert abs(score - res.score) < 0.02
    img.set_dct_form(_lib.DCT_EXACT)
    img.close()


def _uneven_crops(n, seed, lo=256, hi=3072):
    """n crops of uneven size (each side uniform in [lo, hi]) from one seeded
    hi x hi image of the generator."""
    rng = np.random.default_rng(seed)
    master = _rgb(hi, hi, seed)
    out = []
    for _ in range(n):
        w, h = (int(v) for v in rng.integers(lo, hi + 1, 2))
        x0, y0 = int(rng.integers(0, hi - w + 1)), int(rng.integers(0, hi - h + 1))
        out.append(np.ascontiguousarray(master[y0:y0 + h, x0:x0 + w]))
    return out


def test_batched_256_uneven_images_equal_one_at_a_time(ctx):
    """256 images of uneven size (sides uniform in [256, 3072]), targets drawn
    from {70, 85, 90}: one batched quality encode (a launch per stage over
    the whole batch each round) gives every image the rung, probes and
    score its own encode gives, and the same quantized planes (checked on
    every 32nd image)."""
    import jpxl_amd

    rgbs = _uneven_crops(256, 9)
    targets = [float(t) for t in np.random.default_rng(10).choice([70, 85, 90], len(rgbs))]
    imgs = [jpxl_amd.Image(ctx, x, 8) for x in rgbs]
    res = jpxl_amd.batch_encode_quality(imgs, targets, "balanced")
    planes = {k: imgs[k].plan_download().quant for k in range(0, len(imgs), 32)}
    for im in imgs:
        im.close()
    for k, x in enumerate(rgbs):
        one = jpxl_amd.Image(ctx, x, 8)
        r1 = one.encode_quality(targets[k], "balanced")
        assert (res[k].rung, res[k].probes, res[k].met, res[k].status) == (r1.rung, r1.probes, r1.met, r1.status), k
        assert res[k].score == r1.score, k
        if k in planes:
            assert np.array_equal(planes[k], one.plan_download().quant), k
        one.close()


def test_served_mixed_requests_equal_each_served_alone(ctx):
    """Three dozen requests with the serving form's mix (lognormal sizes,
    aspect in [0.5, 2], targets {70, 85, 90}, fast or balanced), arriving a
    Poisson number per server step while others are mid-search: every
    result equals the same request encoded alone."""
    import jpxl_amd

    rng = np.random.default_rng(21)
    n = 36
    reqs = []
    for k in range(n):
        mp = float(np.clip(rng.lognormal(np.log(0.25), 0.8), 0.06, 2.0))
        aspect = float(rng.uniform(0.5, 2.0))
        w = max(16, int(round(np.sqrt(mp * 1e6 * aspect))))
        h = max(16, int(round(mp * 1e6 / w)))
        reqs.append((_rgb(w, h, 100 + k), float(rng.choice([70, 85, 90])), str(rng.choice(["fast", "balanced"]))))
    alone = []
    for x, t, e in reqs:
        im = jpxl_amd.Image(ctx, x, 8)
        r = im.encode_quality(t, e)
        alone.append((r.rung, r.probes, r.score, r.status))
        im.close()
    srv = jpxl_amd.Server(ctx)
    imgs = [jpxl_amd.Image(ctx, x, 8) for x, _, _ in reqs]
    tickets, done, nxt = {}, {}, 0
    arrivals = rng.poisson(1.5, 10 * n)
    step = 0
    while len(done) < n:
        for _ in range(int(arrivals[step % len(arrivals)])):
            if nxt < n:
                tickets[nxt] = srv.submit(imgs[nxt], reqs[nxt][1], reqs[nxt][2])
                nxt += 1
        step += 1
        if not tickets:
            continue
        srv.step()
        for k, t in tickets.items():
            if k not in done:
                r = srv.poll(t)
                if r is not None:
                    done[k] = r
    srv.close()
    for k in range(n):
        r = done[k]
        assert (r.rung, r.probes, r.score, r.status) == alone[k], k
    for im in imgs:
        im.close()


def test_server_admission_queues_within_budget(ctx):
    """Requests given as host samples under a budget of about two requests'
    projected footprint: the server queues the rest (queued > 0 seen), the
    admitted footprint never exceeds the budget unless one request alone
    holds memory, each released request returns its footprint, the
    completed-ticket list names exactly the requests that poll done, and
    every result (and the chosen plan of every other request) equals the
    request encoded alone."""
    import jpxl_amd

    rng = np.random.default_rng(31)
    reqs = []
    for k in range(8):
        w, h = (int(v) for v in rng.integers(200, 900, 2))
        reqs.append((_rgb(w, h, 300 + k), float(rng.choice([70, 85, 90])), str(rng.choice(["fast", "balanced"]))))
    fps = [int(ctx.lib.jpxlg_image_footprint(x.shape[1], x.shape[0])) for x, _, _ in reqs]
    alone = []
    for x, t, e in reqs:
        im = jpxl_amd.Image(ctx, x, 8)
        r = im.encode_quality(t, e)
        alone.append(((r.rung, r.probes, r.score, r.status), im.plan_download().quant))
        im.close()
    budget = int(2.2 * float(np.median(fps)))
    srv = jpxl_amd.Server(ctx)
    srv.set_budget(budget)
    tickets = [srv.submit_srgb(x, t, e) for x, t, e in reqs]
    used, b, queued = srv.admission()
    assert (used, b, queued) == (0, budget, len(reqs))
    done, saw_queue = {}, False
    while len(done) < len(reqs):
        srv.step()
        used, _, queued = srv.admission()
        saw_queue |= queued > 0
        live = [k for k in range(len(reqs)) if k not in done]
        # admission is in arrival order: the queued requests are the last ones
        assert used == sum(fps[k] for k in live[:len(live) - queued])
        assert used <= budget or len(live) - queued == 1
        finished = set(srv.completed())
        for k in live:
            r = srv.poll(tickets[k])
            assert (r is not None) == (tickets[k] in finished), k
            if r is not None:
                done[k] = r
                if k % 2 == 0:
                    assert np.array_equal(srv.plan(tickets[k]).quant, alone[k][1]), k
                srv.release(tickets[k])
    assert saw_queue and srv.completed() == []
    assert srv.admission()[:1] == (0,) and srv.admission()[2] == 0
    srv.close()
    for k in range(len(reqs)):
        r = done[k]
        assert (r.rung, r.probes, r.score, r.status) == alone[k][0], k
