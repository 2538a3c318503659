"""The codestream stage on the CPU (attach_entropy + emit_codestream_with_executor,
jpxl-encode-policy/src/lib.rs:1012, jpxl-encode/src/vardct/write.rs:350):
oracle-quantized plans are handed to jpxlg_emit_codestream in the
QuantizedFrameIr layout and read back by the checker (a restatement of the
reference decoder's reading rules, tests/host/codestream_check.cpp), which
must return exactly the plan that went in.  The checker's reading rules are
pinned by the reference's own known-answer tests (jpxl-entropy/src/prefix.rs,
dist.rs, hybrid.rs) and by the JPEG zigzag for the 8x8 natural order.

No GPU: the codestream stage is host code by design (entropy coding stays on
the CPU)."""
import ctypes

import numpy as np
import pytest

import codestream_binding as cs
import oracle_binding as orc


# ---- the checker's reading rules, pinned ---------------------------------------

def test_prefix_codes_follow_the_rfc_example():
    # prefix.rs:442-456: lengths (2, 1, 3, 3) over bytes 218, 1 read B, A, C, D
    assert cs.prefix_decode([2, 1, 3, 3], bytes([218, 1]), 4) == ([1, 0, 2, 3], 9)


def test_code_length_code_matches_the_printed_table():
    # prefix.rs:463-486: the fixed code of the code-length alphabet, bits in stream order
    table = {0: [0, 0], 1: [1, 1, 1, 0], 2: [1, 1, 0], 3: [0, 1], 4: [1, 0], 5: [1, 1, 1, 1]}
    for sym, bits in table.items():
        data = bytearray((len(bits) + 7) // 8)
        for i, b in enumerate(bits):
            data[i // 8] |= b << (i % 8)
        assert cs.prefix_decode([2, 4, 3, 2, 2, 4], bytes(data), 1) == ([sym], len(bits))


def test_cluster_maps_follow_the_reference_vectors():
    # dist.rs tests: a simple map of 3 contexts from bytes 165, 0; a non-dense map refused
    assert cs.cluster_map(bytes([165, 0]), 3) == ([0, 1, 1], 9)
    with pytest.raises(ValueError):
        cs.cluster_map(bytes([69, 0b10]), 3)
    # one context reads nothing
    assert cs.cluster_map(bytes([0xFF] * 4), 1) == ([0], 0)


def test_hybrid_integers():
    # hybrid.rs: tokens below the split are literal and read nothing
    for t in range(16):
        assert cs.hybrid(4, 0, 0, t) == (t, 0)
    # (4, 2, 0): token 16 + 4 (n - 4) + m, n - 2 extra bits: value = (4 + m) << (n - 2) | extra
    for v in [16, 17, 31, 100, 1000, 65535, 1 << 20]:
        n = v.bit_length() - 1
        m = (v >> (n - 2)) & 3
        tok = 16 + 4 * (n - 4) + m
        extra = v & ((1 << (n - 2)) - 1)
        data = int(extra).to_bytes(8, "little")
        assert cs.hybrid(4, 2, 0, tok, data) == (v, n - 2)


def test_natural_order_8x8_is_the_jpeg_zigzag():
    # ITU T.81 Figure A.6; varblock.rs:1716-1724 pins its first entries 0, 1, 8
    zigzag = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
              6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45,
              38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63]
    assert cs.natural_order(8).tolist() == zigzag
    assert orc.natural_order(8).reshape(-1).tolist() == zigzag
    for side in (16, 32):
        o = cs.natural_order(side)
        assert sorted(o.tolist()) == list(range(side * side))
        assert np.array_equal(o, orc.natural_order(side).reshape(-1))
        c = side // 8
        assert o[: c * c].tolist() == [y * side + x for y in range(c) for x in range(c)]


# ---- round trips ---------------------------------------------------------------

def _plan(w, h, seed, mixed, gs=4000, bits=8):
    from jpxl_amd.synth import synthetic_rgb

    rgb = synthetic_rgb(w, h, seed, bits)
    xyb, _, _ = orc.prepare_frame(rgb, bits)
    bmap = orc.mixed_map(w, h, seed) if mixed else orc.fixed_dct8_map(w, h)
    coeff = orc.forward_frame(xyb, bmap)
    qp = orc.QuantParams(gs, 1, 16, 3, 2, 1, 1.0, 0.0)
    q, lf = orc.quantize_frame(w, h, bmap, coeff, qp)
    return bmap, q, lf


def _frame(w, h, bits=8, gs=4000, lff=(0, 0), gab=0, epf=1, sharp=4, qm=(3, 2)):
    from jpxl_amd import _lib

    return _lib.FrameDesc(width=w, height=h, bits_per_sample=bits, global_scale=gs, quant_lf=16,
                          x_qm_scale=qm[0], b_qm_scale=qm[1], x_factor_lf=lff[0], b_factor_lf=lff[1],
                          gaborish=gab, epf_iters=epf, sharpness=sharp)


def _check(d, w, h, bmap, q, lf, mul=None, cx=None, cb=None, frame=None):
    assert d.info.stage == 7
    assert (d.info.width, d.info.height) == (w, h)
    assert np.array_equal(d.block_map, bmap)
    assert np.array_equal(d.lf, lf)
    assert np.array_equal(d.quant, cs.llf_zeroed(q, bmap))
    m = np.where(bmap > 0, mul if mul is not None else 1, 0)
    assert np.array_equal(d.hf_mul, m)
    if cx is not None:
        assert np.array_equal(d.cfl_x, cx) and np.array_equal(d.cfl_b, cb)
    else:
        assert not d.cfl_x.any() and not d.cfl_b.any()
    if frame is not None:
        i = d.info
        assert (i.bits, i.global_scale, i.quant_lf, i.x_qm_scale, i.b_qm_scale) == (
            frame.bits_per_sample, frame.global_scale, frame.quant_lf, frame.x_qm_scale, frame.b_qm_scale)
        assert (i.gaborish, i.epf_iters) == (frame.gaborish, frame.epf_iters)
        assert (i.x_factor_lf, i.b_factor_lf) == (frame.x_factor_lf, frame.b_factor_lf)
        assert bool(i.lf_corr_default) == (frame.x_factor_lf == 0 and frame.b_factor_lf == 0)
        assert (d.sharpness == frame.sharpness).all()
        assert i.flags == 0x80 and i.is_last == 1 and i.num_passes == 1


@pytest.mark.parametrize("w,h,mixed", [(8, 8, False), (13, 7, False), (203, 141, False), (203, 141, True),
                                       (256, 256, True), (257, 120, True), (600, 520, True)])
def test_oracle_plans_round_trip(w, h, mixed):
    bmap, q, lf = _plan(w, h, 3, mixed)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap)
    f = _frame(w, h)
    stream, sizes = cs.emit(f, arena, lfg, vbs)
    assert stream[:2] == b"\xff\x0a"
    assert sizes.total == len(stream)
    single = ((w + 255) // 256) * ((h + 255) // 256) == 1
    d = cs.decode(stream)
    assert d.info.num_sections == (1 if single else 2 + d.info.num_lf_groups + d.info.num_groups)
    _check(d, w, h, bmap, q, lf, frame=f)


def test_frame_fields_cfl_and_per_varblock_muls_round_trip():
    w, h = 600, 520
    bmap, q, lf = _plan(w, h, 5, True, gs=9000)
    rng = np.random.default_rng(7)
    th, tw = (bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8
    cx = rng.integers(-128, 128, (th, tw)).astype(np.int8)
    cb = rng.integers(-128, 128, (th, tw)).astype(np.int8)
    mul = rng.integers(1, 9, bmap.shape).astype(np.uint32)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap, mul)
    for f in [_frame(w, h, 12, 9000, (-5, 7), 1, 3, 0, (0, 7)), _frame(w, h, 16, 9000, (127, -128), 0, 0, 7),
              _frame(w, h, 8, 9000, (0, 3), 1, 2, 2, (2, 2))]:
        stream, _ = cs.emit(f, arena, lfg, vbs, cx, cb)
        _check(cs.decode(stream), w, h, bmap, q, lf, mul, cx, cb, frame=f)


def test_frames_over_several_lf_groups_round_trip():
    # 2100 x 70: two LF groups side by side (LF groups are 2048 pixels wide)
    w, h = 2100, 70
    bmap, q, lf = _plan(w, h, 9, True)
    rng = np.random.default_rng(2)
    th, tw = (bmap.shape[0] + 7) // 8, (bmap.shape[1] + 7) // 8
    cx = rng.integers(-30, 31, (th, tw)).astype(np.int8)
    cb = rng.integers(-30, 31, (th, tw)).astype(np.int8)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap)
    stream, sizes = cs.emit(_frame(w, h), arena, lfg, vbs, cx, cb)
    d = cs.decode(stream)
    assert d.info.num_lf_groups == 2
    assert sizes.lf_groups > 0 and sizes.hf_global > 0 and sizes.pass_groups > 0
    _check(d, w, h, bmap, q, lf, cx=cx, cb=cb)


def test_large_coefficients_and_empty_plans_round_trip():
    # |q| up to 2^20 (the quantizer's bound) and an all-zero plan
    w, h = 96, 80
    bmap = orc.mixed_map(w, h, 1)
    rng = np.random.default_rng(3)
    q = (rng.integers(-(1 << 20), 1 << 20, (3, 8 * ((h + 7) // 8), 8 * ((w + 7) // 8)))
         * (rng.random((3, 8 * ((h + 7) // 8), 8 * ((w + 7) // 8))) < 0.2)).astype(np.int32)
    lf = rng.integers(-(1 << 15), 1 << 15, (3,) + bmap.shape).astype(np.int32)
    for qq, ll in [(q, lf), (np.zeros_like(q), np.zeros_like(lf))]:
        arena, lfg, vbs = cs.ir_from_footprint(qq, ll, bmap)
        stream, _ = cs.emit(_frame(w, h), arena, lfg, vbs)
        _check(cs.decode(stream), w, h, bmap, qq, ll)


def test_bytes_do_not_depend_on_the_worker_count():
    # results identical for any worker count (EncodeExecutor, jpxl-perceptual/src/executor.rs:16-20)
    w, h = 700, 300
    bmap, q, lf = _plan(w, h, 4, True)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap)
    streams = {cs.emit(_frame(w, h), arena, lfg, vbs, threads=t)[0] for t in (1, 3, 8)}
    assert len(streams) == 1


def test_pricing_sizes_and_a_short_buffer():
    from jpxl_amd import _lib

    w, h = 300, 200
    bmap, q, lf = _plan(w, h, 2, True)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap)
    f = _frame(w, h)
    stream, sizes = cs.emit(f, arena, lfg, vbs)
    assert sizes.total == len(stream)
    assert sizes.image_headers + (sizes.frame_header_bits + sizes.toc_bits) // 8 + sizes.lf_global + \
        sizes.lf_groups + sizes.hf_global + sizes.pass_groups == sizes.total
    lib = _lib.load()
    table = cs.varblock_table(vbs)
    out = np.empty(len(stream) - 1, np.uint8)
    n = ctypes.c_size_t()
    rc = lib.jpxlg_emit_codestream(ctypes.byref(f), cs._p(arena), cs._p(lfg), table, len(vbs), None, None, 2,
                                   cs._p(out), out.size, ctypes.byref(n), None)
    assert rc == _lib.E_TOO_SMALL and n.value == len(stream)


def test_bad_plans_are_refused():
    from jpxl_amd import _lib

    w, h = 64, 64
    bmap, q, lf = _plan(w, h, 2, True)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap)
    lib = _lib.load()
    n = ctypes.c_size_t()

    def rc(vb, f=None):
        t = cs.varblock_table(vb)
        return lib.jpxlg_emit_codestream(ctypes.byref(f or _frame(w, h)), cs._p(arena), cs._p(lfg), t, len(vb),
                                         None, None, 1, None, 0, ctypes.byref(n), None)

    assert rc(vbs) == 0
    assert rc(vbs[:-1]) == _lib.E_INVALID  # a block left uncovered
    assert rc(vbs + vbs[:1]) == _lib.E_INVALID  # overlap
    bad = [dict(v) for v in vbs]
    bad[0]["side"] = 64
    assert rc(bad) == _lib.E_INVALID
    bad = [dict(v) for v in vbs]
    bad[0]["start"] = len(arena)
    assert rc(bad) == _lib.E_INVALID  # coefficients outside the arena
    bad = [dict(v) for v in vbs]
    bad[0]["lf_group"] = 1
    assert rc(bad) == _lib.E_INVALID
    bad = [dict(v) for v in vbs]
    bad[0]["hf_mul"] = 0
    assert rc(bad) == _lib.E_RANGE
    assert rc(vbs, _frame(w, h, sharp=8)) == _lib.E_RANGE
    assert rc(vbs, _frame(w, h, lff=(128, 0))) == _lib.E_RANGE


@pytest.mark.parametrize("w,h,mixed", [(8, 8, False), (203, 141, True), (600, 520, True)])
def test_runs_give_the_same_stream(w, h, mixed):
    """jpxlg_emit_codestream_runs over the runs of an IR writes the bytes
    jpxlg_emit_codestream writes over the IR itself."""
    bmap, q, lf = _plan(w, h, 9, mixed)
    rng = np.random.default_rng(2)
    mul = rng.integers(1, 5, bmap.shape).astype(np.uint32)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap, mul)
    f = _frame(w, h, gab=1, epf=2)
    ref, _ = cs.emit(f, arena, lfg, vbs)
    ra, runs = cs.runs_from_ir(arena, vbs)
    assert ra.size < arena.size // 2  # the runs are the coefficients up to the last non-zero ones
    for threads in (1, 4):
        got, _ = cs.emit_runs(f, ra, runs, lfg, vbs, threads=threads)
        assert got == ref


def test_runs_that_break_their_definition_are_refused():
    from jpxl_amd import _lib

    w, h = 203, 141
    bmap, q, lf = _plan(w, h, 9, True)
    arena, lfg, vbs = cs.ir_from_footprint(q, lf, bmap)
    f = _frame(w, h)
    ra, runs = cs.runs_from_ir(arena, vbs)
    i = next(k for k, r in enumerate(runs) if r[1] >= 3 and r[2] >= 2)
    st, n, nz = runs[i]

    def refused(ra2, runs2):
        with pytest.raises(RuntimeError) as e:
            cs.emit_runs(f, ra2, runs2, lfg, vbs)
        assert e.value.args[0] == _lib.E_INVALID

    bad = list(runs)
    bad[i] = (st, n, nz + 1)  # more non-zeros declared than present
    refused(ra, bad)
    bad[i] = (st, n, nz - 1)  # fewer
    refused(ra, bad)
    r2 = ra.copy()
    r2[st + n - 1] = 0  # the run no longer ends on a non-zero
    refused(r2, runs)
    bad = list(runs)
    bad[i] = (ra.size - 1, n, nz)  # outside the arena
    refused(ra, bad)
    bad = list(runs)
    side = vbs[i // 3]["side"]
    bad[i] = (st, side * side, nz)  # longer than the HF positions
    refused(ra, bad)
    # and the library is unharmed: the good runs still give the IR's stream
    assert cs.emit_runs(f, ra, runs, lfg, vbs)[0] == cs.emit(f, arena, lfg, vbs)[0]
