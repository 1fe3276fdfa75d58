"""The Python wrappers check every device buffer before a launch: device,
dtype, contiguity and size (the C-ABI itself checks only for null pointers
and plane geometry)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def J():
    import jixel_amd
    return jixel_amd


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


@pytest.fixture(scope="module")
def fe(J):
    f = J.FrontEnd(0)
    yield f
    f.close()


W, H = 72, 40


def rgb_dev(torch, ch=3):
    rng = np.random.default_rng(1)
    return torch.from_numpy(rng.integers(0, 256, (H, W, ch), dtype=np.uint8)).cuda()


def rejects(J, fn, text):
    with pytest.raises(J.JxgError) as e:
        fn()
    assert e.value.code == -1 and text in str(e.value), str(e.value)


def test_rgb_checks(J, fe, torch):
    rgb = rgb_dev(torch, 4)
    rejects(J, lambda: fe.xyb(rgb.cpu(), W, H), "expected a tensor on")
    rejects(J, lambda: fe.xyb(rgb.to(torch.int16), W, H), "uint8")
    rejects(J, lambda: fe.xyb(rgb[..., :3], W, H), "pixels must be contiguous")  # 4-byte pixel pitch
    rejects(J, lambda: fe.xyb(rgb[:-1], W, H), "shape")
    fe.xyb(rgb[..., :3].contiguous(), W, H)  # accepted
    # a row stride of its own is fine: a crop of a wider image
    wide = rgb_dev(torch, 3)
    big = torch.zeros((H, W + 8, 3), dtype=torch.uint8, device="cuda")
    big[:, :W] = wide
    a = fe.xyb(big[:, :W], W, H)
    b = fe.xyb(wide, W, H)
    assert torch.equal(a, b)


def test_plane_and_buffer_checks(J, fe, torch):
    xyb = fe.xyb(rgb_dev(torch), W, H)
    p = J.frame_params(J.EncodeConfig())
    nb = ((W + 7) // 8) * ((H + 7) // 8)
    rejects(J, lambda: fe.fill_quant_field(xyb.double(), W, H, 1.0, 1.0), "float32")
    rejects(J, lambda: fe.fill_quant_field(xyb[:, :, :-4], W, H, 1.0, 1.0), "contiguous")
    rejects(J, lambda: fe.fill_quant_field(xyb, W, H, 1.0, 1.0,
                                           out=torch.empty(nb - 1, dtype=torch.uint8, device="cuda")),
            "the stage needs")
    qf = fe.fill_quant_field(xyb, W, H, p.distance, 1.0 / p.scale)
    rejects(J, lambda: fe.select_strategy(xyb, W, H, qf.to(torch.int32), p), "uint8")
    rejects(J, lambda: fe.select_strategy(xyb, W, H, qf.t(), p), "contiguous")
    ytox = torch.zeros(1, dtype=torch.int8, device="cuda")
    rejects(J, lambda: fe.select_strategy(xyb, W, H, qf, p, ytox=ytox), "the stage needs")
    cells = fe.select_strategy(xyb, W, H, qf, p)
    o = fe.outputs(W, H)
    o["levels"][1] = o["levels"][1][:-1]
    rejects(J, lambda: fe.forward_quantize(xyb, W, H, cells, qf, p, outputs=o), "levels[1]")
    o = fe.outputs(W, H)
    o["dc"][2] = o["dc"][2].to(torch.int32)
    rejects(J, lambda: fe.forward_quantize(xyb, W, H, cells, qf, p, outputs=o), "int16")


def test_alpha_checks(J, fe, torch):
    xyb = fe.xyb(rgb_dev(torch), W, H)
    alpha = torch.full((H, W), 255, dtype=torch.uint8, device="cuda")
    rejects(J, lambda: fe.fill_invisible(xyb, W, H, alpha.cpu()), "expected a tensor on")
    rejects(J, lambda: fe.fill_invisible(xyb, W, H, alpha[:, :-1]), "contiguous")
    rejects(J, lambda: fe.fill_invisible(xyb, W, H, alpha.reshape(-1)[:-1]), "the stage needs")
    rejects(J, lambda: fe.alpha_fringe_clear(alpha.float(), W, H), "uint8")
    fe.fill_invisible(xyb, W, H, alpha)


def test_run_checks(J, fe, torch):
    rgb = rgb_dev(torch)
    p = J.frame_params(J.EncodeConfig())
    nb = ((W + 7) // 8) * ((H + 7) // 8)
    rejects(J, lambda: fe.run(rgb, W, H, p, qf=torch.empty(nb - 1, dtype=torch.uint8,
                                                           device="cuda")), "qf")
    rejects(J, lambda: fe.run(rgb.cpu(), W, H, p), "expected a tensor on")
    out = fe.run(rgb, W, H, p)
    assert out["qf"].shape == ((H + 7) // 8, (W + 7) // 8)
