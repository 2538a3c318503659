"""The codestream reader against the reference's own VarDCT fixtures
(JPXL/tests/fixtures/handmade, encoded by cjxl v0.13.0; copied byte for byte
under tests/golden/jxl with the facts of their provenance sidecars in
manifest.json).  The reader (tests/host/codestream_check.cpp, restating the
reference decoder's reading rules) must consume each stream completely: every
section ends exactly at its TOC length, every ANS stream ends in its final
state, every pass group's HF coefficients account for their non-zero counts.
These streams exercise what the library's own streams do not (ANS
distributions, LZ77 parameters, clustered and custom HF block contexts,
global MA trees with the weighted predictor, custom coefficient orders,
DctSelect types up to DCT64x64), so the rules the emitted streams are
checked against are pinned by libjxl's output.  Grading the decoded pixels
against the fixtures' djxl .npy needs the inverse transforms of every
DctSelect type, which this library does not build."""
import hashlib
import json
import os

import numpy as np
import pytest

import codestream_binding as cs

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jxl")
MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))


@pytest.mark.parametrize("fx", MANIFEST["fixtures"], ids=lambda f: f["file"][:2])
def test_reference_fixture_is_read_completely(fx):
    data = open(os.path.join(HERE, fx["file"]), "rb").read()
    assert len(data) == fx["bytes"] and hashlib.sha256(data).hexdigest() == fx["sha256"]
    d = cs.decode(data)
    i = d.info
    assert i.stage == 7
    assert (i.width, i.height) == (fx["width"], fx["height"])
    assert i.colour_space == (1 if fx["grey"] else 0)
    if fx["default_filters"]:
        assert i.gaborish == 1 and 1 <= i.epf_iters <= 3
    else:
        assert i.gaborish == 0 and i.epf_iters == 0
    bw, bh = (i.width + 7) // 8, (i.height + 7) // 8
    assert sum(i.dct_select_histogram) > 0
    # every block is covered exactly once by the varblocks the stream places
    assert not (d.block_map == 255).any()
    assert i.header_bytes + i.toc_sum == len(data)
    # the sharpness planes are valid EPF sharpness values
    assert d.sharpness.max() <= 7


def test_fixtures_exercise_the_rules_the_library_does_not_emit():
    seen = {"global_tree": 0, "custom_orders": 0, "multi_section": 0, "dct_types": set()}
    for fx in MANIFEST["fixtures"]:
        i = cs.decode(open(os.path.join(HERE, fx["file"]), "rb").read()).info
        seen["global_tree"] += i.global_tree
        seen["custom_orders"] += i.used_orders != 0
        seen["multi_section"] += i.num_sections > 1
        seen["dct_types"] |= {k for k, v in enumerate(i.dct_select_histogram) if v}
    assert seen["global_tree"] >= 10 and seen["custom_orders"] >= 2 and seen["multi_section"] >= 1
    assert {0, 2, 4, 5, 6, 10, 11, 18, 19} <= seen["dct_types"]


def test_a_libjxl_plan_re_emitted_reads_back_identically():
    """Fixture 64 (128x2176, two LF groups, every varblock DCT8x8, a custom
    DCT8 coefficient order): the plan the reader returns, written again by
    jpxlg_emit_codestream (natural orders, its own entropy coding), reads
    back to the same plan."""
    from jpxl_amd import _lib

    fx = next(f for f in MANIFEST["fixtures"] if f["file"].startswith("64_"))
    d = cs.decode(open(os.path.join(HERE, fx["file"]), "rb").read())
    i = d.info
    assert i.num_lf_groups == 2 and i.used_orders == 1
    assert set(np.unique(d.block_map)) == {8}
    arena, lfg, vbs = cs.ir_from_footprint(d.quant, d.lf, d.block_map, d.hf_mul)
    f = _lib.FrameDesc(width=i.width, height=i.height, bits_per_sample=i.bits, global_scale=i.global_scale,
                       quant_lf=i.quant_lf, x_qm_scale=i.x_qm_scale, b_qm_scale=i.b_qm_scale,
                       x_factor_lf=i.x_factor_lf, b_factor_lf=i.b_factor_lf, gaborish=i.gaborish,
                       epf_iters=i.epf_iters, sharpness=int(d.sharpness[0, 0]))
    assert (d.sharpness == d.sharpness[0, 0]).all()
    stream, _ = cs.emit(f, arena, lfg, vbs, d.cfl_x, d.cfl_b)
    e = cs.decode(stream)
    assert e.info.stage == 7
    assert np.array_equal(e.quant, d.quant) and np.array_equal(e.lf, d.lf)
    assert np.array_equal(e.block_map, d.block_map) and np.array_equal(e.hf_mul, d.hf_mul)
    assert np.array_equal(e.cfl_x, d.cfl_x) and np.array_equal(e.cfl_b, d.cfl_b)
    assert (e.info.global_scale, e.info.quant_lf) == (i.global_scale, i.quant_lf)
