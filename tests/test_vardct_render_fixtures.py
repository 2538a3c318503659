"""The render graded against the reference's own decodes.

The reference's VarDCT fixtures with a committed djxl decode (fixtures 50-53,
63 of JPXL/tests/fixtures/handmade, copied into tests/golden/jxl/ with their
.npy, sha256 in manifest.json) are read by the reader and reconstructed by
tests/vardct_render.py, whose every step the encoder's render also takes runs
through the oracle's own primitive (dequantization matrices, bias adjustment,
Mul, qm factor, LF multipliers, DCTs, ScaleF, Gaborish, EPF, opsin inverse).
The result must conform to the djxl decode at the reference decoder's own
thresholds (jpxl-decode/tests/e2e_vardct.rs:282-380: 18181-3 Table A.1's
no-filters class, peak 0.004 / channel RMSE 1e-5, and the with-filters
class, peak 0.06 / RMSE 0.02).  It does so with orders of magnitude to spare,
so the tighter bounds below are asserted too.  Planted mistakes in the
reconstruction (LF smoothing dropped, the portrait LLF left untransposed,
the linear-segment reading of the sRGB curve below zero, no Gaborish) fail
these grades; EPF switched off stays inside the loose filtered class but
leaves the margin by two orders of magnitude.

Then the chain to the product: this library's streams, reconstructed the
same way, equal the oracle's encoder-side render of the plan they carry
(the GPU render is bit-exact to that render in the GPU parity tests)."""
import json
import os

import numpy as np
import pytest

import codestream_binding as cs
import oracle_binding as orc
import vardct_render as vr

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jxl")
MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))
GRADED = [f for f in MANIFEST["fixtures"] if "reference_decode" in f]

# (peak, channel RMSE): 18181-3 Table A.1 as the reference decoder's tests apply it
NO_FILTERS = (0.004, 1e-5)
WITH_FILTERS = (0.06, 0.02)


def _grade(fixture):
    stream = open(os.path.join(HERE, fixture["file"]), "rb").read()
    ref = np.load(os.path.join(HERE, fixture["reference_decode"]["file"]))[0]
    ours = vr.render_output(cs.decode_full(stream))
    assert ours.shape == ref.shape
    d = np.abs(ours.astype(np.float64) - ref.astype(np.float64))
    rmse = np.sqrt((d ** 2).reshape(-1, ref.shape[2]).mean(axis=0))
    return float(d.max()), rmse


def test_the_graded_fixtures_are_the_reference_ones():
    import hashlib

    assert [f["file"][:2] for f in GRADED] == ["50", "51", "52", "53", "63"]
    for f in GRADED:
        rd = f["reference_decode"]
        blob = open(os.path.join(HERE, rd["file"]), "rb").read()
        assert hashlib.sha256(blob).hexdigest() == rd["sha256"]
        a = np.load(os.path.join(HERE, rd["file"]))
        assert list(a.shape) == rd["shape"] == [1, f["height"], f["width"], 1 if f["grey"] else 3]


@pytest.mark.parametrize("fixture", GRADED, ids=[f["file"][:2] for f in GRADED])
def test_fixture_conforms_to_the_reference_decode(fixture):
    peak, rmse = _grade(fixture)
    limit_peak, limit_rmse = WITH_FILTERS if fixture["default_filters"] else NO_FILTERS
    assert peak <= limit_peak and (rmse <= limit_rmse).all(), (peak, rmse)
    # the margin actually held: every fixture, filtered or not, inside the
    # strict class by two orders of magnitude on the peak
    assert peak < 5e-5 and (rmse < 5e-6).all(), (peak, rmse)


def test_the_fixtures_cover_the_transform_types_they_grade():
    seen = set()
    for f in GRADED:
        d = cs.decode_full(open(os.path.join(HERE, f["file"]), "rb").read())
        seen |= {v[0] for v in d.varblocks}
        assert not d.info.flags & 0x80  # adaptive LF smoothing runs on every one
    # DCT8, DCT2x2, DCT16, DCT32, DCT16x8, DCT32x16, DCT16x32, DCT64, DCT64x32
    assert seen == {0, 2, 4, 5, 6, 10, 11, 18, 19}


def _planted(monkeypatch, name, fn):
    monkeypatch.setattr(vr, name, fn)
    return [_grade(f) for f in GRADED]


def _fails(grades):
    out = []
    for f, (peak, rmse) in zip(GRADED, grades):
        lp, lr = WITH_FILTERS if f["default_filters"] else NO_FILTERS
        out.append(peak > lp or (rmse > lr).any())
    return out


def test_planted_mistakes_fail_the_grade(monkeypatch):
    # I.5.2's smoothing dropped: the unfiltered grey fixtures fail their class,
    # the filtered ones leave the margin every correct reconstruction holds
    g = _planted(monkeypatch, "adaptive_smoothing", lambda p, m: p.copy())
    assert _fails(g)[:2] == [True, True]
    assert all(peak > 1e-3 for peak, _ in g[2:4])
    monkeypatch.undo()
    # the portrait LLF left in image orientation: the fixtures with portrait
    # rectangles (51-53) fail
    good = vr.llf_from_lf

    def untransposed(ds, rect):
        r, c = vr.BLOCKS[ds]
        return good(ds, rect.T.copy()) if r > c else good(ds, rect)

    assert _fails(_planted(monkeypatch, "llf_from_lf", untransposed))[1:4] == [True] * 3
    monkeypatch.undo()
    # the signed value taking the linear branch of the sRGB curve: the
    # out-of-gamut fixture fails (the reference's flip point, color.rs:440-478)
    lin = lambda v: np.where(v <= 0.0031308, 12.92 * v,  # noqa: E731
                             1.055 * np.power(np.abs(v), 1 / 2.4) - 0.055).astype(np.float32)
    assert _fails(_planted(monkeypatch, "linear_to_srgb", lin))[4]
    monkeypatch.undo()
    # Annex J switched off: the two filtered fixtures fail
    monkeypatch.setattr(orc, "gaborish", lambda p: p)
    assert _fails([_grade(f) for f in GRADED])[2:4] == [True, True]
    monkeypatch.undo()
    # EPF switched off: inside the loose filtered class, far outside the margin
    monkeypatch.setattr(orc, "epf_step", lambda s, p, sg: p)
    g = [_grade(f) for f in GRADED]
    assert all(peak > 4e-3 for peak, _ in g[2:4])


def test_64_point_idct_agrees_with_the_oracle_transform_convention():
    """The f64 inverse used for 64-point sides follows the oracle's f32
    transform (orientation and scale) wherever both apply."""
    rng = np.random.default_rng(3)
    for rows, cols in [(32, 32), (16, 32), (32, 16), (8, 16), (16, 8), (8, 8)]:
        c = rng.standard_normal((min(rows, cols), max(rows, cols))).astype(np.float32)
        a = vr.idct_2d(c, rows, cols)
        m = c.astype(np.float64) if cols > rows else c.astype(np.float64).T
        b = vr._idct_basis(rows) @ m @ vr._idct_basis(cols).T
        assert np.abs(a - b).max() < 2e-5 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("gab,epf,lff", [(0, 0, (0, 0)), (1, 3, (-9, 14)), (1, 1, (5, -3))])
def test_library_streams_reconstruct_to_the_encoder_render(gab, epf, lff):
    """A plan over DCT8/16/32 with CfL tiles and LF factors, written by
    jpxlg_emit_codestream, reconstructed by the fixture-graded decoder
    render, equals the oracle's encoder-side render of the same plan."""
    from jpxl_amd import _lib
    from jpxl_amd.synth import synthetic_rgb

    w, h = 203, 141
    rgb = synthetic_rgb(w, h, 8, 8)
    xyb, _, _ = orc.prepare_frame(rgb, 8)
    bmap = orc.mixed_map(w, h, 8)
    coeff = orc.forward_frame(xyb, bmap)
    gs, mul = 7000, 3
    q, lf = orc.quantize_frame(w, h, bmap, coeff, orc.QuantParams(gs, mul, 16, 3, 2, 1, 1.0, 0.0))
    rng = np.random.default_rng(5)
    th, tw = (bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8
    cx = rng.integers(-40, 40, (th, tw)).astype(np.int8)
    cb = rng.integers(-40, 40, (th, tw)).astype(np.int8)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap, np.full(bmap.shape, mul, np.uint32))
    frame = _lib.FrameDesc(width=w, height=h, bits_per_sample=8, global_scale=gs, quant_lf=16, x_qm_scale=3,
                           b_qm_scale=2, x_factor_lf=lff[0], b_factor_lf=lff[1], gaborish=gab, epf_iters=epf,
                           sharpness=5)
    stream, _ = cs.emit(frame, arena, lfg, vbs, cx, cb)
    full = cs.decode_full(stream)
    assert full.info.flags & 0x80  # this library's frames skip the LF smoothing
    ours = vr.render_xyb(full)
    rp = orc.RenderParams(gs, mul, 16, 3, 2, epf, gab, 5, 8)
    _, enc = orc.render(w, h, bmap, q, lf, rp, cx, cb, lff)
    assert np.abs(ours - enc).max() < 2e-6, np.abs(ours - enc).max()
