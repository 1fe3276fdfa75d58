"""GPU parity of the chroma-HF protection of the quant field (stage and
front-end flag) against the oracle, bit for bit."""
import numpy as np
import pytest

from test_gpu_parity import first_mismatch, same_bits, to_dev

pytestmark = pytest.mark.gpu


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


def chromatic_rgb(h, w, seed):
    """Saturated hue edges over mild luma: the blocks the protection raises."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    hue = ((x // 3 + y // 5) % 3)
    base = np.stack([(hue == 0) * 220, (hue == 1) * 200, (hue == 2) * 230], axis=2) + 15
    noise = rng.integers(-10, 11, (h, w, 3))
    flat = (x // 32 + y // 32) % 4 == 0  # some flat gray areas too
    img = np.where(flat[..., None], 128 + noise, base + noise)
    return np.clip(img, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h,d", [(64, 64, 0.8), (131, 77, 0.5), (200, 150, 1.0),
                                   (97, 33, 1.5), (9, 1, 0.9), (1, 1, 0.9), (120, 80, 2.35)])
def test_chroma_hf_bit_exact(J, fe, torch, oracle, w, h, d, fused):
    print("chroma_hf", w, h, d, fused, flush=True)
    rgb = chromatic_rgb(h, w, w * h)
    xyb = fe.xyb(to_dev(torch, rgb), w, h)
    src = xyb[:, :h, :w].cpu().numpy()
    rng = np.random.default_rng(3)
    qf = rng.integers(1, 256, ((h + 7) // 8, (w + 7) // 8)).astype(np.uint8)
    oracle.set_fused(fused)
    want = oracle.chroma_hf_protection(src, qf, d)
    oracle.set_fused(True)
    d_q = to_dev(torch, qf)
    fe.chroma_hf_protection(xyb, w, h, d_q, d, fused)
    got = d_q.cpu().numpy()
    assert np.array_equal(got, want), first_mismatch(got, want)
    if 0.4 < d < 2.0 and w * h > 64:
        assert (want != qf).any()  # the case exercises the raise


@pytest.mark.parametrize("fused", [True, False])
def test_front_end_chroma_hf_bit_exact(J, fe, torch, oracle, fused):
    w, h = 160, 96
    rgb = chromatic_rgb(h, w, 7)
    p = J.frame_params(J.EncodeConfig(fused=fused).with_distance(0.8))
    p.flags = p.flags | J.FLAG_CHROMA_HF
    want = oracle.front_end(rgb, p, oracle.tables_for(p), threads=4)
    o = fe.run(to_dev(torch, rgb), w, h, p)
    for k in ("qf", "cells", "ytox", "ytob"):
        assert same_bits(o[k].cpu().numpy(), want[k]), k
    for c in range(3):
        assert same_bits(o["levels"][c].cpu().numpy(), want["levels"][c]), c
        assert same_bits(o["dc"][c].cpu().numpy(), want["dc"][c]), c
