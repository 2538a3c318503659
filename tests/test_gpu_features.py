"""The quality search's first-guess features on the GPU, through the C ABI:
jpxlg_image_quality_features against the oracle's source features, 32x32
preanalysis and DCT8 summary (quality_features.rs, preanalysis.rs,
lib.rs:2769-2955), and the first probe of jpxlg_encode_quality against the
routing of quality.rs:2485-2530 restated over the oracle's features."""
import math

import numpy as np
import pytest

import cpu_encode
import oracle_binding as orc

pytestmark = pytest.mark.gpu

# (width, height, seed, bits, grayscale)
CASES = [(96, 80, 1, 8, False), (257, 131, 3, 16, False), (640, 480, 4, 8, False), (40, 31, 5, 8, False),
         (300, 200, 6, 8, True), (1531, 1027, 7, 12, False), (2048, 2048, 2, 16, False)]


@pytest.fixture(scope="module")
def ctx():
    import jpxl_amd

    c = jpxl_amd.Context(0)
    yield c
    c.close()


def _rgb(w, h, seed, bits, gray):
    from jpxl_amd.synth import synthetic_rgb

    x = synthetic_rgb(w, h, seed, 16 if bits > 8 else 8)
    if bits not in (8, 16):
        x = (x >> (16 - bits)).astype(np.uint16)
    if gray:
        x = np.repeat(x[..., 1:2], 3, axis=2)
    return np.ascontiguousarray(x)


def _fields(s):
    out = {}
    for name, _ in type(s)._fields_:
        v = getattr(s, name)
        out[name] = list(v) if name == "share" else v
    return out


@pytest.mark.parametrize("w,h,seed,bits,gray", CASES)
def test_quality_features_match_the_oracle(ctx, w, h, seed, bits, gray):
    """Source features and preanalysis bit-exact; the DCT8 summary's counts,
    histogram quantiles and block count exact, its energy sums (summed per
    block and per workgroup run instead of as one running sum) within
    1e-12 relative."""
    import jpxl_amd

    print(f"features {w}x{h} seed {seed} bits {bits} gray {gray}", flush=True)
    rgb = _rgb(w, h, seed, bits, gray)
    img = jpxl_amd.Image(ctx, rgb, bits)
    sf, ts = img.quality_features()
    img.close()
    xyb, _, g = orc.prepare_frame(rgb, bits)
    assert g == gray
    osf, ots = cpu_encode.oracle_features(rgb, bits, xyb, g)
    assert _fields(sf) == _fields(osf)
    assert sf.has_preanalysis == (w >= 32 and h >= 32)
    a, b = _fields(ts), _fields(ots)
    for k in ("blocks", "ln_ac_y_q50", "ln_ac_y_q90", "ln_ac_y_q99", "near_zero_frac_1e3", "near_zero_frac_1e2"):
        assert a[k] == b[k], k
    for k in ("ln_ac_y_mean", "high_low_ratio", "directional_asymmetry", "chroma_ac_ratio"):
        assert math.isclose(a[k], b[k], rel_tol=1e-12, abs_tol=1e-300), (k, a[k], b[k])
    assert math.isclose(a["dc_variance_y"], b["dc_variance_y"], rel_tol=1e-9, abs_tol=1e-15)


def _routed_start(sf, ts, target, reserve, legacy):
    """quality.rs:2485-2530 in the default build (one-shot-controller,
    flagged-median-start, case-predictor)."""
    import jpxl_amd
    from jpxl_amd import _lib

    p = jpxl_amd.predict_crossing(sf, ts, target, reserve)
    if p is not None:
        if p.fallback == _lib.FALLBACK_NONE:
            return p.candidate_rung, p.local_loss_exponent
        if p.ood_flags == 0:
            return p.median_rung, p.local_loss_exponent
        if p.large_frame_only_ood:
            return p.median_rung, None
    return legacy, None


@pytest.mark.parametrize("w,h,seed,bits,target,effort", [
    (640, 480, 4, 8, 85.0, "balanced"), (640, 480, 4, 8, 70.0, "fast"), (257, 131, 3, 16, 90.0, "balanced"),
    (96, 80, 1, 8, 85.0, "fast"), (2048, 2048, 2, 16, 85.0, "balanced")])
def test_first_probe_follows_the_crossing_predictor(ctx, w, h, seed, bits, target, effort):
    import jpxl_amd

    print(f"start {w}x{h} seed {seed} q{target} {effort}", flush=True)
    rgb = _rgb(w, h, seed, bits, False)
    img = jpxl_amd.Image(ctx, rgb, bits)
    res = img.encode_quality(target, effort)
    img.close()
    xyb, _, g = orc.prepare_frame(rgb, bits)
    osf, ots = cpu_encode.oracle_features(rgb, bits, xyb, g)
    reserve = 0.06 if effort == "fast" else 0.03
    legacy = cpu_encode.predicted_rung(xyb, target)
    want, _ = _routed_start(osf, ots, target, reserve, legacy)
    assert res.predicted_rung == want
    assert res.probe_rungs[0] == want
