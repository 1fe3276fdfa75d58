"""GPU parity of the Fast / Fastest strategy decision with the coefficient
re-tiling (layout_dp::retile_coefficient, src/ac_strategy/selection/
layout_dp.rs:2381): the square-merge search, the aligned and (at Fast) the
(4, 4)-shifted re-tiling passes and adjust_quant_field, against the
oracle's restatement, bit for bit (cells and quant field), on photo-like
and smooth content, ragged sizes and several DC groups in each direction.
Each case also shows that the re-tiling acted: the square search alone
never places a DCT16X8 / DCT8X16 or an unaligned DCT16X16."""
import numpy as np
import pytest

from test_gpu_parity import _stage_inputs, first_mismatch, padded, same_bits, to_dev

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def J():
    import jixel_amd
    return jixel_amd


@pytest.fixture(scope="module")
def fe(J):
    f = J.FrontEnd(0)
    yield f
    f.close()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


def retile_marks(J, cells):
    """Counts of pieces only the re-tiling places: pairs and unaligned
    16x16 squares."""
    S = J.STRATEGY
    s = cells >> 1
    first = (cells & 1) == 1
    pairs = int(np.sum(first & ((s == S["DCT16X8"]) | (s == S["DCT8X16"]))))
    ys, xs = np.nonzero(first & (s == S["DCT16X16"]))
    unaligned = int(np.sum(((xs % 2) != 0) | ((ys % 2) != 0)))
    return pairs, unaligned


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("speed", ["fast", "fastest"])
@pytest.mark.parametrize("w,h,distance,seed", [(256, 256, 2.35, 3), (300, 260, 1.8, 4),
                                               (2100, 72, 2.35, 5), (72, 2100, 3.0, 6),
                                               (517, 389, 4.6, 7), (130, 94, 2.0, 8)])
def test_fast_decision_with_retiling_bit_exact(J, fe, torch, oracle, w, h, distance, seed, speed,
                                               fused):
    log("retile select", w, h, distance, speed, fused)
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, distance, seed, fused, cmap=True)
    p.speed = int(J.Speed[speed.upper()])
    tables = oracle.tables_for(p)
    cells_w, qf_w = oracle.select_strategy(xyb, qf, ytox, ytob, tables, p)
    oracle.set_fused(True)
    qf_d = to_dev(torch, qf)
    cells = fe.select_strategy(to_dev(torch, padded(J, xyb)), w, h, qf_d, p, to_dev(torch, ytox),
                               to_dev(torch, ytob)).cpu().numpy()
    assert same_bits(cells, cells_w), first_mismatch(cells, cells_w)
    assert same_bits(qf_d.cpu().numpy(), qf_w), first_mismatch(qf_d.cpu().numpy(), qf_w)
    pairs, unaligned = retile_marks(J, cells_w)
    log("  pairs", pairs, "unaligned 16x16", unaligned)
    assert pairs > 0
    if speed == "fastest":
        assert unaligned == 0  # layout_dp_unaligned is off at Fastest


@pytest.mark.parametrize("fused", [True, False])
def test_fast_decision_smooth_content_bit_exact(J, fe, torch, oracle, fused):
    from jixel_amd.synth import smooth_rgba
    w, h = 640, 480
    rgba = smooth_rgba(w, h, seed=9)
    rgb = np.ascontiguousarray(rgba[..., :3])
    oracle.set_fused(fused)
    xyb = oracle.xyb_image(rgb, J.XYB_SPEC_FWD)
    p = J.frame_params(J.EncodeConfig(fused=fused).with_quality(75.0))
    qf = oracle.fill_quant_field(xyb, p.distance, 1.0 / p.scale)
    tiles = ((h + 63) // 64, (w + 63) // 64)
    ytox = np.zeros(tiles, np.int8)
    ytob = np.zeros(tiles, np.int8)
    tables = oracle.tables_for(p)
    cells_w, qf_w = oracle.select_strategy(xyb, qf, ytox, ytob, tables, p)
    oracle.set_fused(True)
    qf_d = to_dev(torch, qf)
    cells = fe.select_strategy(to_dev(torch, padded(J, xyb)), w, h, qf_d, p, to_dev(torch, ytox),
                               to_dev(torch, ytob)).cpu().numpy()
    assert same_bits(cells, cells_w), first_mismatch(cells, cells_w)
    assert same_bits(qf_d.cpu().numpy(), qf_w), first_mismatch(qf_d.cpu().numpy(), qf_w)
    log("  strategies", np.bincount(cells_w.reshape(-1) >> 1, minlength=5)[:5],
        "marks", retile_marks(J, cells_w))
