"""The quality search's crossing predictor on the CPU (host code of the
library, no device): the embedded tables against values printed in the
reference's generated sources, the reference's own predictor tests
(jpxl-encode-policy/src/quality_prediction.rs:708-938) restated, and the
library's predict_cases / predict_v2 against an independent numpy
restatement on perturbed feature vectors."""
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLES = os.path.join(ROOT, "jpxl_amd", "csrc", "predictor_tables.bin")


@pytest.fixture(scope="module")
def jl():
    import jpxl_amd
    from jpxl_amd import _lib

    try:
        _lib.load()
    except (RuntimeError, OSError) as e:  # the library is built by build()
        pytest.skip(f"libjpxlg.so not loadable: {e}")
    return jpxl_amd


# ---- the tables -------------------------------------------------------------


def read_tables():
    b = open(TABLES, "rb").read()
    assert b[:8] == b"JPXLQPT1"
    d2, k, d, n, t, nb, _ = struct.unpack("<7I", b[8:36])
    v = np.frombuffer(b[36:], "<f8")
    o = 0

    def take(m):
        nonlocal o
        r = v[o:o + m]
        o += m
        return r

    tab = {"d2": d2, "K": k, "d": d, "N": n, "T": t, "neighbours": nb}
    tab["v2_centers"], tab["v2_scales"] = take(d2), take(d2)
    tab["v2_zrange"] = take(2 * d2).reshape(d2, 2)
    tab["knots"] = take(k * (1 + 6 * (d2 + 1))).reshape(k, -1)
    tab["ratio"], tab["spread"], tab["eps"] = take(3)
    tab["weights"], tab["centers"], tab["scales"] = take(d), take(d), take(d)
    tab["knot_targets"] = take(t)
    tab["zrange"] = take(2 * d).reshape(d, 2)
    tab["cases"] = take(n * (d + 2 * t)).reshape(n, -1)
    assert o == v.size
    return tab


def test_tables_hold_the_generated_models():
    """Counts, constants and spot values as printed in
    quality_predictor_v2.rs and quality_predictor_cases.rs."""
    tab = read_tables()
    assert (tab["d2"], tab["K"], tab["d"], tab["N"], tab["T"], tab["neighbours"]) == (19, 7, 44, 1203, 7, 3)
    assert (tab["ratio"], tab["spread"], tab["eps"]) == (3.0, 0.9, 0.05)
    assert list(tab["knot_targets"]) == [30.0, 50.0, 70.0, 80.0, 85.0, 90.0, 95.0]
    assert list(tab["knots"][:, 0]) == [30.0, 50.0, 70.0, 80.0, 85.0, 90.0, 95.0]
    # QPV2_KNOTS[0].lower[0..2] (quality_predictor_v2.rs:117-119)
    assert tab["knots"][0, 1] == 8.114267187113095 and tab["knots"][0, 2] == -0.08228926450496746
    # QPV2_CASES_FEATURE_WEIGHTS[0..12], QPV2_CASES_FEATURE_CENTERS[0], [6]
    assert list(tab["weights"][:13]) == [1.0, 1.0, 1.0, 2.0, 0.0, 1.0, 4.0, 0.0, 1.0, 0.0, 1.0, 0.5, 4.0]
    assert tab["centers"][0] == -9.256104977766855 and tab["centers"][6] == 23.12777507146957
    # QPV2_CASES_FEATURE_Z_RANGE[0] and [43]
    assert tuple(tab["zrange"][0]) == (-6.258251183075504, 3.6636579287505864)
    assert tuple(tab["zrange"][43]) == (-3.604369970842619, 3.6852615373742506)
    # the last case's ln_crossing and ln_beta[0]
    last = tab["cases"][-1]
    assert list(last[44:51]) == [7.950585613027057, 8.68331638705579, 9.561452297929677, 10.139925699606547,
                                 10.56471568055867, 11.163439067887658, 12.294763438059356]
    assert last[51] == -0.6085772031152448


# ---- the reference's fixture (quality_prediction.rs:712-771) -----------------


def fixture(jl):
    from jpxl_amd import _lib

    sf = _lib.SourceFeatures(width=640, height=480, grayscale=0, luma_variance_q10=1e-6, luma_variance_q50=1e-4,
                             luma_variance_q90=1e-2, chroma_variance_q50=1e-5, flat_fraction=0.1, edge_proxy=1e-2,
                             has_preanalysis=1, cells=48705)
    for i, v in enumerate([0.0451904, 0.724505, 0.0380659, 0.0, 0.108305, 0.0675085, 0.0164254]):
        sf.share[i] = v
    for n, v in dict(log2var_mean=7.07718, log2var_std=1.73362, log2var_q10=5.12928, log2var_q90=8.98868,
                     edge_mean=0.0716309, edge_q90=0.230469, flat_mean=0.722714, noise_mean=0.0984498,
                     noise_q90=0.0, chroma_log2_mean=5.71906, chroma_log2_q90=8.65464, orient_entropy_mean=0.428383,
                     dir_dominance_mean=0.739334, axis_aligned_mean=0.759818, hetero_frac=0.186819).items():
        setattr(sf, n, v)
    ts = _lib.TransformSummary(blocks=4800, ln_ac_y_mean=-5.0, ln_ac_y_q50=-6.0, ln_ac_y_q90=-3.0,
                               ln_ac_y_q99=-2.0, high_low_ratio=0.2, directional_asymmetry=0.3, chroma_ac_ratio=0.15,
                               near_zero_frac_1e3=0.6, near_zero_frac_1e2=0.8, dc_variance_y=0.02)
    return sf, ts


def _tup(p):
    return tuple(getattr(p, f) for f, _ in type(p)._fields_)


def _eff(jl, rung):
    q = jl.rung_quantizer(int(rung))
    return q.global_scale * q.hf_mul


def test_knot_model_is_monotone_in_the_target(jl):
    """predictions_are_monotone_in_the_target (quality_prediction.rs:786)."""
    sf, ts = fixture(jl)
    prev = 0
    for target in [30.0, 50.0, 70.0, 80.0, 85.0, 90.0, 95.0]:
        p = jl.predict_crossing(sf, ts, target, 0.03, 2)
        s = _eff(jl, p.candidate_rung)
        assert s >= prev, target
        prev = s


def test_large_frame_flag_names_size_as_the_only_ood_signal(jl):
    """the_large_frame_flag_names_size_as_the_only_ood_signal (:808)."""
    from jpxl_amd import _lib

    sf, ts = fixture(jl)
    p = jl.predict_crossing(sf, ts, 85.0, 0.03, 2)
    assert not p.large_frame_only_ood or p.ood_flags == 1 << 6
    sf.width = sf.height = 1 << 15
    p = jl.predict_crossing(sf, ts, 85.0, 0.03, 2)
    assert p.ood_flags & (1 << 6)
    assert bool(p.large_frame_only_ood) == (p.ood_flags == 1 << 6)
    sf.width = sf.height = 16
    p = jl.predict_crossing(sf, ts, 85.0, 0.03, 2)
    assert not p.large_frame_only_ood and p.ood_flags & _lib.OOD_TINY_FRAME


def test_case_predictor_is_monotone_and_aims_finer_than_its_median(jl):
    """the_case_predictor_is_monotone_and_aims_finer_than_its_median (:840)."""
    from jpxl_amd import _lib

    sf, ts = fixture(jl)
    prev = 0
    for target in [30.0, 50.0, 70.0, 80.0, 85.0, 90.0, 95.0]:
        p = jl.predict_crossing(sf, ts, target, 0.03, 1)
        assert p.median_rung >= prev
        assert p.candidate_rung >= p.median_rung
        assert p.interval_low <= p.median_rung <= p.interval_high
        if p.fallback == _lib.FALLBACK_NONE:
            assert p.ood_flags == 0
        else:
            assert p.fallback == _lib.FALLBACK_WIDE_INTERVAL and p.ood_flags == _lib.OOD_CASE_SPREAD
        assert 0.2 <= p.local_loss_exponent <= 3.0
        prev = p.median_rung
    balanced = jl.predict_crossing(sf, ts, 85.0, 0.03, 1)
    fast = jl.predict_crossing(sf, ts, 85.0, 0.06, 1)
    assert fast.candidate_rung >= balanced.candidate_rung and fast.median_rung == balanced.median_rung
    sf.width = sf.height = 64
    p = jl.predict_crossing(sf, ts, 85.0, 0.03, 1)
    assert p.ood_flags & _lib.OOD_TINY_FRAME and p.fallback == _lib.FALLBACK_OOD_FEATURE
    assert not p.large_frame_only_ood
    sf.width = sf.height = 1 << 15
    p = jl.predict_crossing(sf, ts, 85.0, 0.03, 1)
    assert p.ood_flags == 1 << 6 and p.large_frame_only_ood
    sf, ts = fixture(jl)
    cases = jl.predict_crossing(sf, ts, 85.0, 0.03, 1)
    both = jl.predict_crossing(sf, ts, 85.0, 0.03, 0)
    if cases.fallback == _lib.FALLBACK_NONE:
        assert _tup(both) == _tup(cases)
    else:
        assert _tup(both) == _tup(jl.predict_crossing(sf, ts, 85.0, 0.03, 2))
    sf.has_preanalysis = 0
    assert _tup(jl.predict_crossing(sf, ts, 85.0, 0.03, 0)) == _tup(jl.predict_crossing(sf, ts, 85.0, 0.03, 2))
    assert jl.predict_crossing(sf, ts, 85.0, 0.03, 1) is None


# ---- an independent restatement of the predictors ----------------------------


def _loss(s):
    return max(100.0 - s, 1e-3)


def _st_vector(sf, ts):
    def ln_eps(v):
        return np.log(max(float(np.float32(v)), 0.0) + 1e-9)

    w, h = float(max(sf.width, 1)), float(max(sf.height, 1))
    return np.array([ln_eps(sf.luma_variance_q10), ln_eps(sf.luma_variance_q50), ln_eps(sf.luma_variance_q90),
                     ln_eps(sf.chroma_variance_q50), float(np.float32(sf.flat_fraction)), ln_eps(sf.edge_proxy),
                     np.log2(w * h), np.log(w / h), 1.0 if sf.grayscale else 0.0, ts.chroma_ac_ratio,
                     ts.dc_variance_y, ts.directional_asymmetry, ts.high_low_ratio, ts.ln_ac_y_mean, ts.ln_ac_y_q50,
                     ts.ln_ac_y_q90, ts.ln_ac_y_q99, ts.near_zero_frac_1e2, ts.near_zero_frac_1e3])


def _pa_vector(sf):
    f = [float(np.float32(x)) for x in list(sf.share)] + [float(np.float32(getattr(sf, n))) for n in (
        "log2var_mean", "log2var_std", "log2var_q10", "log2var_q90", "edge_mean", "edge_q90", "flat_mean",
        "noise_mean", "noise_q90", "chroma_log2_mean", "chroma_log2_q90", "orient_entropy_mean",
        "dir_dominance_mean", "axis_aligned_mean", "hetero_frac")]
    return np.array(f + [np.log(f[11] + 1e-4), np.log(1.0 - f[0] + 1e-3), np.log(f[14] + 0.1)])


def _interp_knots(xs, vals, x):
    if x >= xs[0]:
        return vals[0]
    if x <= xs[-1]:
        return vals[-1]
    for i in range(len(xs) - 1):
        if xs[i + 1] <= x <= xs[i]:
            return vals[i] + (xs[i] - x) / (xs[i] - xs[i + 1]) * (vals[i + 1] - vals[i])
    return vals[-1]


def np_predict_cases(tab, sf, ts, target, reserve):
    raw = np.concatenate([_st_vector(sf, ts), _pa_vector(sf)])
    z = (raw - tab["centers"]) / tab["scales"]
    cz = tab["cases"][:, :44]
    d = np.sqrt(np.sum((tab["weights"] * (cz - z)) ** 2, axis=1))
    order = sorted(range(len(d)), key=lambda i: (d[i], i))[:3]
    order = [i for i in order if d[i] <= d[order[0]] * 3.0 + 1e-9]
    xs = [np.log(_loss(t)) for t in tab["knot_targets"]]
    xt = np.log(_loss(target))
    xa = np.log(_loss(100.0 - _loss(target) * (1.0 - reserve)))
    tot = med = cand = beta = 0.0
    lo, hi = np.inf, -np.inf
    for i in order:
        w = 1.0 / (d[i] + 0.05)
        c = tab["cases"][i]
        at_t = _interp_knots(xs, c[44:51], xt)
        med += w * at_t
        cand += w * _interp_knots(xs, c[44:51], xa)
        beta += w * _interp_knots(xs, c[51:58], xt)
        lo, hi = min(lo, at_t), max(hi, at_t)
        tot += w
    med /= tot
    return med, max(cand / tot, med), lo, hi, beta / tot


def np_predict_v2(tab, sf, ts, target):
    z = (_st_vector(sf, ts) - tab["v2_centers"]) / tab["v2_scales"]
    k = tab["knots"]
    W = 20
    ev = []
    for row in k:
        e = [np.log(_loss(row[0]))] + [row[1 + f * W] + float(np.dot(row[2 + f * W:1 + (f + 1) * W], z))
                                       for f in range(6)]
        ev.append(e)
    ev = np.array(ev)
    for q in range(1, 5):
        ev[:, q] = np.maximum.accumulate(ev[:, q])
    x = np.log(_loss(target))
    out = []
    for fld in range(1, 7):
        xs = list(ev[:, 0])
        out.append(_interp_knots(xs, list(ev[:, fld]), x))
    return out  # lower, median, candidate, upper, beta, saturation logit


def test_library_predictors_match_a_numpy_restatement(jl):
    """Perturbed copies of the fixture: the library's rungs bracket the
    numpy restatement's crossings (median and interval ends rounded down,
    candidate up), and the loss exponents agree."""
    tab = read_tables()
    rng = np.random.default_rng(5)
    checked = 0
    for trial in range(60):
        sf, ts = fixture(jl)
        sf.width, sf.height = int(rng.integers(200, 3000)), int(rng.integers(200, 3000))
        for n in ("luma_variance_q10", "luma_variance_q50", "luma_variance_q90", "chroma_variance_q50", "edge_proxy"):
            setattr(sf, n, getattr(sf, n) * float(np.exp(rng.normal(0, 0.7))))
        sf.flat_fraction = float(rng.uniform(0, 0.6))
        for n in ("log2var_mean", "edge_mean", "noise_mean", "chroma_log2_mean", "hetero_frac"):
            setattr(sf, n, getattr(sf, n) * float(np.exp(rng.normal(0, 0.3))))
        for n in ("ln_ac_y_mean", "ln_ac_y_q50", "ln_ac_y_q90", "ln_ac_y_q99"):
            setattr(ts, n, getattr(ts, n) + float(rng.normal(0, 0.5)))
        ts.high_low_ratio *= float(np.exp(rng.normal(0, 0.4)))
        target = float(rng.choice([55.0, 70.0, 85.0, 90.0]))
        reserve = float(rng.choice([0.03, 0.06]))
        p = jl.predict_crossing(sf, ts, target, reserve, 1)
        med, cand, lo, hi, beta = np_predict_cases(tab, sf, ts, target, reserve)
        for rung, v, up in ((p.median_rung, med, False), (p.interval_low, lo, False), (p.candidate_rung, cand, True),
                            (p.interval_high, hi, True)):
            s = float(np.exp(v))
            e = _eff(jl, rung)
            if up:
                assert e >= s * (1 - 1e-12) and (rung == 0 or _eff(jl, rung - 1) < s * (1 + 1e-12)), trial
            else:
                assert e <= s * (1 + 1e-12) and _eff(jl, rung + 1) > s * (1 - 1e-12), trial
        assert abs(p.local_loss_exponent - min(max(np.exp(beta), 0.2), 3.0)) < 1e-12
        q = jl.predict_crossing(sf, ts, target, reserve, 2)
        lower, median, candidate, upper, lb, sat = np_predict_v2(tab, sf, ts, target)
        for rung, v, up in ((q.median_rung, median, False), (q.candidate_rung, candidate, True)):
            s = float(np.exp(v))
            if not np.isfinite(s) or s < 1.0 or s > 4e6:
                continue
            e = _eff(jl, rung)
            if up:
                assert e >= s * (1 - 1e-12), trial
            else:
                assert e <= s * (1 + 1e-12) and _eff(jl, rung + 1) > s * (1 - 1e-12), trial
        assert abs(q.local_loss_exponent - min(max(np.exp(lb), 0.2), 3.0)) < 1e-12
        assert abs(q.saturation_risk - 1.0 / (1.0 + np.exp(-min(max(sat, -30.0), 30.0)))) < 1e-12
        checked += 1
    assert checked == 60


def test_first_probe_routing(jl):
    """The library's routing (quality.rs:2485-2530, default build) against
    its restatement: the candidate when the prediction is confident, the
    median (with the exponent) when only uncertain, the median without it
    for a frame that is only too large, else the legacy rung."""
    import ctypes

    import cpu_encode
    from jpxl_amd import _lib

    cc = cpu_encode.lib()
    rng = np.random.default_rng(11)
    seen = set()
    for trial in range(80):
        sf, ts = fixture(jl)
        sf.width, sf.height = (int(v) for v in rng.choice([64, 300, 1500, 40000], 2))
        sf.has_preanalysis = int(rng.integers(0, 2))
        sf.luma_variance_q50 *= float(np.exp(rng.normal(0, 1.5)))
        ts.ln_ac_y_mean += float(rng.normal(0, 2.0))
        target, effort = float(rng.choice([50.0, 85.0, 95.0])), int(rng.integers(0, 2))
        reserve = 0.06 if effort == 0 else 0.03
        legacy = 4321
        p = jl.predict_crossing(sf, ts, target, reserve)
        if p.fallback == _lib.FALLBACK_NONE:
            want, beta, route = p.candidate_rung, p.local_loss_exponent, "candidate"
        elif p.ood_flags == 0:
            want, beta, route = p.median_rung, p.local_loss_exponent, "median"
        elif p.large_frame_only_ood:
            want, beta, route = p.median_rung, None, "large"
        else:
            want, beta, route = legacy, None, "legacy"
        seen.add(route)
        r, b = ctypes.c_uint32(), ctypes.c_double()
        cc.cc_quality_start(ctypes.byref(sf), ctypes.byref(ts), target, effort, legacy, ctypes.byref(r),
                            ctypes.byref(b))
        assert r.value == want, (trial, route)
        assert (beta is None and np.isnan(b.value)) or b.value == beta, (trial, route)
    assert {"candidate", "legacy", "large"} <= seen
