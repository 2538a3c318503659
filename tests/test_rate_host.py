"""The byte-rate controller's ladder search compiled host-only (g++) and
driven on the CPU against the reference's own rate-loop tests
(jpxl-encode-policy/src/rate.rs:2465-2915), restated over closed-form size
functions of the quantizer (global_scale x HfMul): landing under and near a
target, aimed steps, a spike above the crossing, a non-monotone pocket
against the brute-force optimum over the whole ladder, the fill's hidden
notch, the hard price cap, the floor refusal and saturation."""
import ctypes

import numpy as np
import pytest

PRICE = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64))
UNREACHABLE, BUDGET = 1, 2


@pytest.fixture(scope="module")
def cc():
    import cpu_encode

    lib = ctypes.CDLL(cpu_encode.build())
    u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
    lib.cc_rate_search.restype = ctypes.c_int
    lib.cc_rate_search.argtypes = [u32, u64, u64, ctypes.c_double, u32, u32, PRICE, ctypes.POINTER(u32),
                                   ctypes.POINTER(u64), ctypes.POINTER(i32), ctypes.POINTER(u64), ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.POINTER(i32)]
    lib.cc_rate_cold_step.restype = u32
    lib.cc_rate_cold_step.argtypes = [u32, ctypes.c_int]
    lib.cc_rate_interpolated.restype = ctypes.c_int
    lib.cc_rate_interpolated.argtypes = [u32, u64, u32, u64, u64, ctypes.POINTER(u32)]
    lib.cc_ladder_fields.restype = u32
    lib.cc_ladder_fields.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.cc_effective_scale.restype = u64
    lib.cc_effective_scale.argtypes = [u32]
    lib.cc_top.restype = u32
    gs = np.zeros(1 << 20, np.uint32)
    mul = np.zeros(1 << 20, np.uint32)
    n = lib.cc_ladder_fields(gs.ctypes.data, mul.ctypes.data)
    lib.gs, lib.mul, lib.n = gs[:n].astype(np.int64), mul[:n].astype(np.int64), n
    return lib


def search(cc, target, size_of, start=None, tol=(8, 0.01), max_prices=40, fill=4):
    """search_ladder(start, quant_lf 16, target, tolerance, budget, price)."""
    calls = []

    def price(rung, out):
        calls.append(rung)
        out[0] = int(size_of(int(cc.gs[rung]), int(cc.mul[rung]), rung))
        return 0

    cb = PRICE(price)
    rung, nbytes, sat, floor, nt = (ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_int32(), ctypes.c_uint64(),
                                    ctypes.c_int32())
    tr = np.zeros(256, np.uint32)
    tb = np.zeros(256, np.uint64)
    start = start if start is not None else 32768 - 1  # Rung::for_global_scale(32768)
    rc = cc.cc_rate_search(start, target, tol[0], tol[1], max_prices, fill, cb, ctypes.byref(rung),
                           ctypes.byref(nbytes), ctypes.byref(sat), ctypes.byref(floor), tr.ctypes.data,
                           tb.ctypes.data, ctypes.byref(nt))
    trace = list(zip(tr[: nt.value].tolist(), tb[: nt.value].tolist()))
    return rc, rung.value, nbytes.value, bool(sat.value), floor.value, trace, calls


def smooth(gs, mul, _r):  # rate.rs:2547-2550
    return 100 + gs * mul * 3


def slack_of(target, tol=(8, 0.01)):
    return max(tol[0], int(np.floor(tol[1] * target)))


def test_cold_bracketing_is_geometric_in_effective_scale(cc):
    below = 65536 - 1
    assert cc.cc_effective_scale(below) == 65536
    doubled = cc.cc_rate_cold_step(below, 1)
    assert cc.cc_effective_scale(doubled) == 131072
    assert cc.cc_rate_cold_step(doubled, 0) == below
    assert cc.cc_effective_scale(below * 2 + 1) > cc.cc_effective_scale(doubled)
    assert cc.cc_rate_cold_step(0, 0) == 0
    assert cc.cc_rate_cold_step(cc.cc_top(), 1) == cc.cc_top()


@pytest.mark.parametrize("target", [200, 1000, 50000, 98404])
def test_the_loop_lands_under_a_target_and_close_to_it(cc, target):
    rc, rung, nbytes, _, _, trace, calls = search(cc, target, smooth)
    assert rc == 0 and nbytes <= target
    assert target - nbytes <= max(slack_of(target), 3)
    assert len(trace) <= 40
    if target != 200:  # aimed_stepping_reaches_the_target_in_fewer_prices
        assert len(calls) <= 16


def test_an_already_priced_rung_is_priced_once(cc):
    _, _, _, _, _, trace, calls = search(cc, 50000, smooth)
    assert len(calls) == len(set(calls)) == len(trace)


def test_aimed_steps(cc):
    out = ctypes.c_uint32()
    assert cc.cc_rate_interpolated(100, 1000, 100000, 31_600_000, 2000, ctypes.byref(out))
    mid = 100 + (100000 - 100) // 2
    assert 100 < out.value < 100000 and out.value < mid // 10
    assert not cc.cc_rate_interpolated(10, 100, 11, 200, 150, ctypes.byref(out))
    assert not cc.cc_rate_interpolated(10, 500, 1000, 400, 450, ctypes.byref(out))
    assert not cc.cc_rate_interpolated(10, 0, 1000, 400, 200, ctypes.byref(out))


def _optimum(cc, size_of, target):
    sizes = np.asarray(size_of(cc.gs, cc.mul, np.arange(cc.n)), np.int64)  # every rung of the ladder
    feas = sizes[sizes <= target]
    return int(feas.max())


def test_a_spike_above_the_crossing_does_not_fence_off_finer_rungs(cc):
    target = 10000

    def price(gs, mul, r):
        scale = gs * mul
        base = 100 + scale * 3
        return np.where((scale >= 3000) & (scale < 3200), base + 50000, base)

    rc, _, nbytes, _, _, _, _ = search(cc, target, price, tol=(0, 0.0))
    assert rc == 0 and nbytes <= target
    opt = _optimum(cc, price, target)
    assert nbytes == opt and opt > 3200 * 3


def test_the_loop_survives_a_non_monotone_pocket(cc):
    target = 10000

    def price(gs, mul, r):
        scale = gs * mul
        base = 100 + scale * 3
        pocket = (scale >= 2000) & (scale < 4000)
        return np.where(pocket, np.where(scale % 2 == 0, np.maximum(base - 4000, 0), base + 4000), base)

    rc, _, nbytes, _, _, trace, _ = search(cc, target, price)
    assert rc == 0 and nbytes <= target and len(trace) <= 40
    assert nbytes == max(b for _, b in trace if b <= target)
    # the search walked into the pocket: a finer rung priced smaller
    by_rung = sorted(trace)
    assert any(b2 < b1 for (r1, b1), (r2, b2) in zip(by_rung, by_rung[1:]))
    assert nbytes == _optimum(cc, price, target)


def test_the_fill_recovers_a_notch_behind_an_infeasible_one(cc):
    def price(gs, mul, r):
        gs = np.asarray(gs)
        return np.where(gs <= 5000, 9000, np.where(gs == 5001, 10001, np.where(gs == 5002, 9500, 20000)))

    rc, rung, nbytes, _, _, _, _ = search(cc, 10000, price, tol=(0, 0.0))
    assert rc == 0 and nbytes == 9500 and rung == 5001


def test_the_price_budget_is_a_hard_cap(cc):
    rc, _, nbytes, _, _, trace, calls = search(cc, 10000, smooth, tol=(0, 0.0), max_prices=6)
    assert rc == 0 and len(calls) == len(trace) <= 6 and nbytes <= 10000


def test_a_target_below_the_floor_is_refused_with_the_floor(cc):
    rc, _, _, _, floor, _, _ = search(cc, 50, smooth)
    assert rc == UNREACHABLE and floor == 103


def test_a_target_above_the_ceiling_saturates_at_the_finest_rung(cc):
    rc, rung, _, sat, _, _, _ = search(cc, (1 << 62), smooth)
    assert rc == 0 and sat and rung == cc.cc_top()


def test_saturation_reports_a_feasible_top_even_when_a_coarser_pocket_wins(cc):
    top = cc.cc_top()
    rc, rung, nbytes, sat, _, trace, _ = search(cc, 1000, lambda g, m, r: np.where(np.asarray(r) == top, 1, 900),
                                                start=0, tol=(0, 0.0))
    assert rc == 0 and rung < top and nbytes == 900 and sat
    assert any(r == top and b <= 1000 for r, b in trace)
