"""GPU parity of the Fast decision at d <= 1.709, where the SSIM rerank
(find_rerank_downgrades, src/ac_strategy/selection.rs:2718-3232) splits
merged transforms between the re-tiling and adjust_quant_field: cells and
field bit-exact against the oracle on photo-like and smooth content, ragged
sizes, two DC groups across, both fmla modes, and with the CVVDP arbiter
(whose zero rerank margin splits nothing)."""
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


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h,distance,seed,arbiter", [(256, 256, 1.6, 3, 0), (300, 260, 1.2, 4, 0),
                                                       (2100, 72, 1.7, 5, 0), (517, 389, 1.0, 7, 0),
                                                       (130, 94, 0.6, 8, 0), (256, 256, 1.3, 9, 1)])
def test_fast_decision_with_rerank_bit_exact(J, fe, torch, oracle, w, h, distance, seed, arbiter,
                                             fused):
    log("rerank select", w, h, distance, arbiter, fused)
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, distance, seed, fused, cmap=True)
    p.speed = int(J.Speed.FAST)
    p.arbiter = arbiter
    tables = oracle.tables_for(p)
    cells_w, qf_w = oracle.select_strategy(xyb, qf, ytox, ytob, tables, p)
    oracle.set_fused(True)
    qf_d = to_dev(torch, qf)
    cells = fe.select_strategy(to_dev(torch, padded(J, xyb)), w, h, qf_d, p, to_dev(torch, ytox),
                               to_dev(torch, ytob)).cpu().numpy()
    assert same_bits(cells, cells_w), first_mismatch(cells, cells_w)
    assert same_bits(qf_d.cpu().numpy(), qf_w), first_mismatch(qf_d.cpu().numpy(), qf_w)
    # the rerank acted: without it the oracle keeps more merges (none are
    # split with the CVVDP arbiter's zero margin)
    oracle.rerank_enable(False)
    try:
        oracle.set_fused(fused)
        cells_n, _ = oracle.select_strategy(xyb, qf, ytox, ytob, tables, p)
    finally:
        oracle.rerank_enable(True)
        oracle.set_fused(True)
    split = int(np.sum(cells_n != cells_w))
    log("  strategies", np.bincount(cells_w.reshape(-1) >> 1, minlength=5)[:5], "cells split", split)
    if arbiter == 0:
        assert split > 0
    else:
        assert split == 0


@pytest.mark.parametrize("fused", [True, False])
def test_fast_decision_with_rerank_smooth_content(J, fe, torch, oracle, fused):
    from jixel_amd.synth import smooth_rgba
    w, h = 640, 480
    rgba = smooth_rgba(w, h, seed=9)
    rgb = np.ascontiguousarray(rgba[..., :3])
    oracle.set_fused(fused)
    xyb = oracle.xyb_image(rgb, J.XYB_SPEC_FWD)
    p = J.frame_params(J.EncodeConfig(fused=fused).with_quality(88.0))
    assert p.distance <= 1.709
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
