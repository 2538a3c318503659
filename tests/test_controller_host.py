"""The library's quality controller and metric pooling compiled host-only
(g++) and driven on the CPU: the reference's controller tests
(jpxl-encode-policy/src/quality.rs:2870-3450) restated against closed-form
score curves, the first-guess predictor's table (quality_predictor.rs,
quality.rs:1021-1079) against the reference's own table, and the metric's
108 weights as the library's host code applies them."""
import ctypes
import math

import numpy as np
import pytest

import reference_restated as rr


MET, ADJ, CAP, SAT_TOP, SAT_FLOOR, UNDER, RESCUED = range(7)
FAST, BALANCED = 0, 1


class Outcome(ctypes.Structure):
    _fields_ = [("rung", ctypes.c_uint32), ("score", ctypes.c_double), ("met", ctypes.c_int32),
                ("status", ctypes.c_int32), ("chosen_fresh", ctypes.c_int32), ("anchor_rung", ctypes.c_uint32),
                ("pixel_probes", ctypes.c_int32), ("structural_builds", ctypes.c_int32),
                ("priced", ctypes.c_int32), ("nprobes", ctypes.c_int32),
                ("probe_rungs", ctypes.c_uint32 * 32), ("probe_scores", ctypes.c_double * 32),
                ("probe_fresh", ctypes.c_int32 * 32)]


@pytest.fixture(scope="module")
def cc():
    import cpu_encode

    lib = ctypes.CDLL(cpu_encode.build())  # the host-only build of the navigator
    d, u32, i32 = ctypes.c_double, ctypes.c_uint32, ctypes.c_int
    lib.cc_met_band.restype = d
    lib.cc_met_band.argtypes = [d, d]
    lib.cc_aim.restype = d
    lib.cc_aim.argtypes = [d, d]
    lib.cc_log_loss_crossing.restype = u32
    lib.cc_log_loss_crossing.argtypes = [u32, d, u32, d, d, ctypes.POINTER(i32)]
    lib.cc_geometric_step.restype = u32
    lib.cc_geometric_step.argtypes = [u32, i32]
    lib.cc_top.restype = u32
    lib.cc_effective_scale.restype = ctypes.c_uint64
    lib.cc_effective_scale.argtypes = [u32]
    lib.cc_rung_for_scale.restype = u32
    lib.cc_rung_for_scale.argtypes = [d]
    lib.cc_predicted_scale.restype = d
    lib.cc_predicted_scale.argtypes = [ctypes.c_float, ctypes.c_float, d]
    lib.cc_features.argtypes = [ctypes.c_void_p, i32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    lib.cc_navigate_curve.restype = i32
    lib.cc_navigate_curve.argtypes = [d, i32, i32, i32, i32, d, u32, d, d, d, i32, ctypes.POINTER(Outcome)]
    lib.cc_pool_terms.restype = d
    lib.cc_pool_terms.argtypes = [ctypes.c_void_p, i32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.POINTER(d)]
    return lib


def _curve_score(cc, rung, a=60.0, b=0.8):
    return 100.0 - a * (cc.cc_effective_scale(rung) / 1000.0) ** -b


def navigate(cc, target, effort=BALANCED, predicted=None, budget=None, a=60.0, b=0.8, fresh_gain=0.0, mode=0,
             max_probes=0):
    o = Outcome()
    if predicted is None:
        # the legacy predictor's start for a mid-texture frame
        predicted = cc.cc_rung_for_scale(cc.cc_predicted_scale(1e-4, 0.1, target))
    if budget is None:
        rc = cc.cc_navigate_curve(target, effort, max_probes, 0, 0, 0.0, predicted, a, b, fresh_gain, mode,
                                  ctypes.byref(o))
    else:
        rc = cc.cc_navigate_curve(target, -1, *budget, predicted, a, b, fresh_gain, mode, ctypes.byref(o))
    assert rc == 0
    return o


def test_the_crossing_stays_strictly_inside_the_bracket(cc):
    ok = ctypes.c_int()
    r = cc.cc_log_loss_crossing(999, 60.0, 7999, 91.0, 85.0, ctypes.byref(ok))
    assert ok.value and 999 < r < 7999
    cc.cc_log_loss_crossing(999, 70.0, 7999, 65.0, 68.0, ctypes.byref(ok))
    assert not ok.value  # a finer point that scored lower cannot be interpolated
    cc.cc_log_loss_crossing(10, 50.0, 11, 60.0, 55.0, ctypes.byref(ok))
    assert not ok.value  # adjacent rungs leave no room


def test_geometric_steps_move_and_stop_at_the_ladder_ends(cc):
    assert cc.cc_geometric_step(999, 1) > 999 and cc.cc_geometric_step(999, 0) < 999
    assert cc.cc_geometric_step(0, 0) == 0
    top = cc.cc_top()
    assert cc.cc_geometric_step(top, 1) == top


def test_the_accept_band_contains_the_crossing_aim(cc):
    for reserve in (0.06, 0.03, 0.02):
        for target in [5.0, 30.0, 50.0, 70.0, 80.0, 85.0, 90.0, 95.0, 99.0]:
            band = cc.cc_met_band(target, reserve)
            assert band >= 1.0
            assert cc.cc_aim(target, reserve) - target < band
            assert abs(band - max(2.0 * reserve * (100.0 - target), 1.0)) < 1e-12
    assert abs(cc.cc_met_band(85.0, 0.03) - 1.0) < 1e-12
    assert abs(cc.cc_met_band(80.0, 0.03) - 1.2) < 1e-12
    assert abs(cc.cc_met_band(50.0, 0.03) - 3.0) < 1e-12
    assert abs(cc.cc_met_band(30.0, 0.03) - 4.2) < 1e-12
    assert abs(cc.cc_met_band(95.0, 0.06) - 1.0) < 1e-12
    assert abs(cc.cc_met_band(85.0, 0.06) - 1.8) < 1e-12


@pytest.mark.parametrize("effort", [FAST, BALANCED])
@pytest.mark.parametrize("target", [50.0, 70.0, 85.0])
def test_every_effort_meets_its_target_inside_its_budget(cc, effort, target):
    o = navigate(cc, target, effort)
    budget = {FAST: (3, 2), BALANCED: (5, 3)}[effort]
    assert o.score >= target and o.met
    assert o.pixel_probes <= budget[0]
    assert o.priced <= budget[1]
    assert o.structural_builds <= 2
    assert o.status in (MET, ADJ, CAP, RESCUED)
    # the chosen score is the curve's score of the chosen rung
    assert abs(o.score - _curve_score(cc, o.rung)) < 1e-9


def test_a_first_probe_inside_the_band_ends_the_search(cc):
    """trusted-first-probe: a feasible first probe within the accept band is
    the answer, one probe (quality.rs:1832-1845)."""
    target = 85.0
    # the rung whose curve score sits half a point above the target
    r = next(r for r in range(1000, 60000) if _curve_score(cc, r) >= target + 0.5)
    o = navigate(cc, target, BALANCED, predicted=r)
    assert o.nprobes == 1 and o.rung == r and o.status == MET


def test_a_near_miss_is_corrected_directly(cc):
    """corrected-stop: a first probe just below the target (implied move
    within the 0.5 gate) is followed by one direct correction that lands in
    the band (quality.rs:1536-1600)."""
    target = 85.0
    r = next(r for r in range(1000, 60000) if _curve_score(cc, r) >= target - 1.0)
    o = navigate(cc, target, BALANCED, predicted=r)
    assert o.nprobes == 2 and o.met and o.status == MET
    assert o.score - target <= cc.cc_met_band(target, 0.03)


def test_a_far_first_probe_brackets_and_tightens(cc):
    target = 85.0
    o = navigate(cc, target, BALANCED, predicted=200)  # far too coarse: implied move above the gate
    assert o.nprobes >= 3 and o.met
    scores = [o.probe_scores[i] for i in range(o.nprobes)]
    assert any(s < target for s in scores) and any(s >= target for s in scores)


def test_running_out_of_probes_reports_under_target_work_cap(cc):
    o = navigate(cc, 99.9, budget=(1, 1, 1, 0.03))
    assert o.status == UNDER and not o.met and o.score < 99.9
    assert o.pixel_probes == 2  # the navigation cap plus the one rescue probe
    assert o.priced == 1


def test_an_unreachable_target_saturates_explicitly(cc):
    o = navigate(cc, 99.99, FAST)
    assert o.status == SAT_TOP and o.score < 99.99
    assert o.rung == cc.cc_top()


def test_a_higher_target_never_costs_a_lower_score(cc):
    lo = navigate(cc, 60.0)
    hi = navigate(cc, 90.0)
    assert hi.score >= lo.score and hi.rung >= lo.rung


def test_a_far_finalist_is_rebuilt_and_rescued(cc):
    """A feasible finalist far from the anchor (> 1.8x in effective scale) is
    re-probed with its own structure; when that still meets the target it is
    the chosen plan (RescuedFreshStructure)."""
    target = 85.0
    o = navigate(cc, target, BALANCED, predicted=200, fresh_gain=0.1)
    fresh = [o.probe_fresh[i] for i in range(o.nprobes)]
    assert fresh[0] == 1
    if o.structural_builds == 2:
        assert sum(fresh) == 2 and o.chosen_fresh and o.status == RESCUED
        assert o.rung != o.anchor_rung
    assert o.met


def test_bisection_mode_meets_the_target(cc):
    o = navigate(cc, 85.0, mode=1, max_probes=8, predicted=0)
    assert o.met and o.pixel_probes <= 8 and o.structural_builds == 1


# ---------------------------------------------------------------------------
# The legacy first-guess predictor (quality.rs:1021-1079, the table of
# quality_predictor.rs)
# ---------------------------------------------------------------------------

def _predict_ref(q50, flat, target, tab):
    """quality.rs:1030-1079 over the reference's table (from the fixture)."""
    t = tab["initial_rung_table"]
    lb = sum(1 for e in t["luma_edges"] if np.float32(q50) >= np.float32(e))
    fb = sum(1 for e in t["flat_edges"] if np.float32(flat) >= np.float32(e))
    tl = math.log(max(100.0 - target, 1e-3))
    ent = sorted([(math.log(max(100.0 - e[0], 1e-3)), math.log(max(e[3], 1))) for e in t["entries"]
                  if e[1] == lb and e[2] == fb], key=lambda v: v[0])
    if len(ent) >= 2:
        seg = next(((a, b) for a, b in zip(ent, ent[1:]) if a[0] <= tl <= b[0]), None)
        if seg is None:
            seg = (ent[0], ent[1]) if tl < ent[0][0] else (ent[-2], ent[-1])
        a, b = seg
        if abs(b[0] - a[0]) > 2.220446049250313e-16:
            return math.exp(a[1] + (tl - a[0]) / (b[0] - a[0]) * (b[1] - a[1]))
    if ent:
        return math.exp(ent[0][1])
    f = t["fallback_log_fit"]
    return math.exp(f[0] + f[1] * tl + f[2] * math.log(float(np.float32(q50)) + 1e-6) + f[3] * float(np.float32(flat)))


def test_predictor_table_equals_the_reference(cc):
    tab = rr.tables()
    edges = tab["initial_rung_table"]["luma_edges"]
    probes = [0.0, 1e-8] + [e for e in edges] + [np.nextafter(np.float32(e), np.float32(0)) for e in edges] + \
        [2e-4, 5e-3]
    for q50 in probes:
        for flat in [0.0, 0.1, 0.2, 0.45, 0.6, 0.9]:
            for target in [20.0, 30.0, 45.0, 70.0, 84.0, 85.0, 91.0, 95.0, 99.0]:
                got = cc.cc_predicted_scale(float(q50), flat, target)
                want = _predict_ref(q50, flat, target, tab)
                assert got == pytest.approx(want, rel=1e-12, abs=0), (q50, flat, target)


def test_the_fallback_prediction_is_monotone_in_the_target(cc):
    a, b, c = (cc.cc_predicted_scale(1e-4, 0.1, t) for t in (50.0, 85.0, 95.0))
    assert a <= b <= c and a < c


def test_source_features_are_floor_indexed(cc):
    v = np.array([5.0, 1.0, 4.0, 2.0, 3.0], np.float32)
    q50, flat = ctypes.c_float(), ctypes.c_float()
    cc.cc_features(v.ctypes.data, len(v), ctypes.byref(q50), ctypes.byref(flat))
    assert q50.value == 3.0 and flat.value == 0.0
    v = np.array([0.0] * 6 + [1e-3] * 4, np.float32)
    cc.cc_features(v.ctypes.data, len(v), ctypes.byref(q50), ctypes.byref(flat))
    assert q50.value == 0.0 and flat.value == pytest.approx(0.6)


# ---------------------------------------------------------------------------
# The library's metric weights (metric_host.cpp) against the reference's
# ---------------------------------------------------------------------------

def test_library_weights_sit_at_their_terms(cc):
    wts = rr.tables()["ssimulacra2"]["weights"]
    for ns in (6, 3):
        sw = np.ones(6, np.int32)
        sh = np.ones(6, np.int32)
        for s in range(ns):
            for c in range(3):
                for k in range(6):
                    sums = np.zeros(18 * 6)
                    m, norm = k % 3, k // 3
                    sums[s * 18 + c * 6 + 2 * m + norm] = 16.0  # 1-norm 16 or 4-norm 2
                    terms = np.zeros(108)
                    raw = ctypes.c_double()
                    cc.cc_pool_terms(sums.ctypes.data, ns, sw.ctypes.data, sh.ctypes.data, terms.ctypes.data,
                                     ctypes.byref(raw))
                    idx = c * ns * 6 + s * 6 + k
                    val = 16.0 if norm == 0 else 2.0
                    assert terms[idx] == val
                    assert raw.value == wts[idx] * val


def test_cpu_encode_runs_the_whole_search():
    """The CPU restatement of the whole encode (the bench's CPU baseline): the
    navigator over oracle probes meets the target on a small image."""
    import cpu_encode
    from jpxl_amd.synth import synthetic_rgb

    rgb = synthetic_rgb(96, 80, 3, 8)
    out, plan = cpu_encode.encode(rgb, 8, 85.0, BALANCED)
    assert out.met and out.score >= 85.0 and 1 <= out.nprobes <= 6
    assert out.probe_fresh[0] == 1
    quant, lf, bmap = plan[:3]
    assert quant.shape == (3, 80, 96) and lf.shape == (3, 10, 12) and bmap.shape == (10, 12)


PROBE_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_double))
PRICE_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64))


def _navigate_priced(cc, target, bytes_of, effort=BALANCED, predicted=None, a=60.0, b=0.8):
    cc.cc_navigate_priced.restype = ctypes.c_int
    cc.cc_navigate_priced.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, PROBE_CB,
                                      PRICE_CB, ctypes.POINTER(Outcome), ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_uint64)]
    priced = []

    def probe(rung, fresh, out):
        out[0] = 100.0 - a * (cc.cc_effective_scale(rung) / 1000.0) ** (-b)
        return 0

    def price(rung, fresh, out):
        priced.append(rung)
        out[0] = bytes_of(rung)
        return 0

    pcb, qcb = PROBE_CB(probe), PRICE_CB(price) if bytes_of else ctypes.cast(None, PRICE_CB)
    o, n, nb = Outcome(), ctypes.c_int32(), ctypes.c_uint64()
    pred = predicted if predicted is not None else cc.cc_rung_for_scale(8000.0)
    assert cc.cc_navigate_priced(target, effort, 0, pred, pcb, qcb, ctypes.byref(o), ctypes.byref(n),
                                 ctypes.byref(nb)) == 0
    return o, n.value, nb.value, priced


def test_priced_finalists_keep_the_smallest_stream(cc):
    """quality.rs:1886-2010: the two coarsest feasible probes are priced
    exactly and the smallest feasible stream wins (the first of equal sizes);
    a single feasible finalist is taken without pricing."""
    seen_two = 0
    for target in np.linspace(60.0, 92.0, 33):
        for pocket in (False, True):
            def bytes_of(rung, pocket=pocket):  # size rising with the rung, with a pocket at odd rungs
                return 1000 + 10 * rung - (25 if pocket and rung % 2 else 0)

            o, n, nb, priced = _navigate_priced(cc, float(target), bytes_of)
            feas = sorted(o.probe_rungs[i] for i in range(o.nprobes) if o.probe_scores[i] >= target)
            fin = sorted(set(feas))[:2]
            if len(fin) < 2:
                assert n == 0 and not priced
                continue
            seen_two += 1
            assert n == 2 and sorted(priced) == fin
            sizes = [bytes_of(r) for r in fin]
            assert o.rung == fin[sizes.index(min(sizes))] and nb == min(sizes)
            # without a price the coarsest feasible rung is chosen
            o2, n2, _, _ = _navigate_priced(cc, float(target), None)
            assert n2 == 0 and o2.rung == fin[0]
    assert seen_two > 0
