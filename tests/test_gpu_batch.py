"""The batched front end (one launch per stage for many images, HIP graph
replay, tile-shape autotuner): every image's outputs bit-identical to the
single-image path, which the other tests hold to the oracle, plus an oracle
check of two batch members directly."""
import numpy as np
import pytest

from test_gpu_parity import first_mismatch, same_bits, to_dev

pytestmark = pytest.mark.gpu

KEYS = ("qf", "cells", "ytox", "ytob")


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


def batch_images(J, torch, n, seed, fused=True):
    """Uneven sizes (not multiples of 8), qualities 50..95, photo content,
    mixed speed tiers."""
    from jixel_amd import synth
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        w, h = int(rng.integers(9, 300)), int(rng.integers(9, 300))
        if i == 0:
            w, h = 517, 263  # one larger image, several tiles and both stage maps
        q = float(rng.uniform(50.0, 95.0))
        img = synth.photo_like(w, h, 1000 + i)
        if i % 5 == 4:
            img = np.ascontiguousarray(img[:, :, 1:2])  # a gray member
        # mixed tiers: Fast (re-tiling, and the quant-field refinement at
        # d >= 1.5), Fastest (aligned re-tiling only) and Ultrafast (the fixed
        # DCT8 layout)
        speed = (J.Speed.FAST, J.Speed.FASTEST, J.Speed.FAST, J.Speed.ULTRAFAST)[i % 4]
        p = J.frame_params(J.EncodeConfig(fused=fused, speed=speed).with_quality(q))
        out.append((img, p))
    return out


def same_outputs(got, want, what):
    for k in KEYS:
        a, b = got[k].cpu().numpy(), want[k].cpu().numpy()
        assert same_bits(a, b), (what, k, first_mismatch(a, b))
    for c in range(3):
        for k in ("levels", "coeffs", "dc"):
            a, b = got[k][c].cpu().numpy(), want[k][c].cpu().numpy()
            assert same_bits(a, b), (what, k, c, first_mismatch(a, b))


@pytest.mark.parametrize("fused", [True, False])
def test_batch_matches_single_image_path(J, fe, torch, fused):
    imgs = batch_images(J, torch, 13, 7 + fused, fused)
    dev = [(to_dev(torch, im), p) for im, p in imgs]
    b = J.Batch(fe, dev)
    try:
        b.run()
        torch.cuda.synchronize()
        # tile-crossing transforms (the shifted re-tiling grid places them)
        # are queued, none is unsupported
        assert b.status()[1] == 0
        for i, (rgb, p) in enumerate(dev):
            want = fe.run(rgb, rgb.shape[1], rgb.shape[0], p)
            same_outputs(b.outputs(i), want, f"image {i} {tuple(rgb.shape)}")
    finally:
        b.close()


def test_batch_members_match_oracle(J, fe, torch, oracle):
    imgs = batch_images(J, torch, 6, 3)
    dev = [(to_dev(torch, im), p) for im, p in imgs]
    b = J.Batch(fe, dev)
    try:
        b.launch()
        torch.cuda.synchronize()
        for i in (0, 4):  # the large colour image and a gray member
            im, p = imgs[i]
            want = oracle.front_end(im, p, oracle.tables_for(p), threads=4)
            got = b.outputs(i)
            for k in KEYS:
                assert same_bits(got[k].cpu().numpy(), want[k]), (i, k)
            for c in range(3):
                assert same_bits(got["levels"][c].cpu().numpy(), want["levels"][c]), (i, c)
                assert same_bits(got["dc"][c].cpu().numpy(), want["dc"][c]), (i, c)
    finally:
        b.close()


def same_dev(a, b):
    """Bit equality of two device tensors (floats compared as their bits)."""
    import torch
    if a.dtype == torch.float32 and b.dtype == torch.float32:
        a, b = a.view(torch.int32), b.view(torch.int32)
    return a.shape == b.shape and a.dtype == b.dtype and bool((a == b).all())


def test_batch_recipe_matches_single_image_path(J, fe, torch, oracle):
    """The first 32 images of the timed batch recipe (sides uniform in
    64..2048, not multiples of 8, quality 50..95, photo-like content with
    seeds 1000 + i, Fast tier): after the autotuner has picked the tile
    shapes, every member equals the single-image path bit for bit, and the
    largest member equals the oracle."""
    from jixel_amd import synth
    dev = torch.device("cuda", 0)
    sizes = synth.batch_sizes(32, 4)  # the 512-image recipe's first 32 draws
    assert max(max(w, h) for w, h, _ in sizes) > 1536
    field = synth.one_over_f_field(4096, 4, dev)
    images = [(synth.photo_like_device(w, h, 1000 + i, dev, field),
               J.frame_params(J.EncodeConfig().with_quality(q)))
              for i, (w, h, q) in enumerate(sizes)]
    del field
    b = J.Batch(fe, images)
    try:
        b.autotune(reps=1)
        b.run()
        torch.cuda.synchronize()
        assert b.status()[1] == 0
        for i, (rgb, p) in enumerate(images):
            got, want = b.outputs(i), fe.run(rgb, rgb.shape[1], rgb.shape[0], p)
            for k in KEYS:
                assert same_dev(got[k], want[k]), (i, sizes[i], k)
            for c in range(3):
                for k in ("levels", "coeffs", "dc"):
                    assert same_dev(got[k][c], want[k][c]), (i, sizes[i], k, c)
        big = max(range(len(sizes)), key=lambda i: sizes[i][0] * sizes[i][1])
        rgb, p = images[big]
        im = rgb.cpu().numpy()
        want = oracle.front_end(im, p, oracle.tables_for(p), threads=8)
        got = b.outputs(big)
        for k in KEYS:
            assert same_bits(got[k].cpu().numpy(), want[k]), (big, sizes[big], k)
        for c in range(3):
            assert same_bits(got["levels"][c].cpu().numpy(), want["levels"][c]), (big, c)
            assert same_bits(got["dc"][c].cpu().numpy(), want["dc"][c]), (big, c)
    finally:
        b.close()


def test_batch_graph_replays_and_autotune(J, fe, torch):
    """The captured graph, replayed several times with eager work and
    synchronizations in between, and after the autotuner changes the tile
    shapes, gives the eager batch's results every time."""
    imgs = batch_images(J, torch, 9, 11)
    dev = [(to_dev(torch, im), p) for im, p in imgs]
    b = J.Batch(fe, dev)
    try:
        b.run()
        torch.cuda.synchronize()
        ref = [{k: (v.clone() if not isinstance(v, list) else [x.clone() for x in v])
                for k, v in b.outputs(i).items()} for i in range(len(dev))]
        for rep in range(3):
            for o in b.outs:  # scribble over the outputs first
                o["qf"].fill_(7)
                o["levels"][1].fill_(-1)
            torch.cuda.synchronize()
            b.launch()
            torch.cuda.synchronize()
            for i in range(len(dev)):
                same_outputs(b.outputs(i), ref[i], f"replay {rep} image {i}")
            _ = (torch.ones(1000, device="cuda") * 2).sum().item()  # eager work in between
        report = b.autotune(reps=2)
        print(report, flush=True)
        assert report.count("\n") == 18
        cfg = b.config()
        assert cfg["aq_strip"] in (16, 32, 64) and cfg["cm_rows"] in (16, 32)
        for o in b.outs:
            o["qf"].fill_(7)
        b.launch()
        torch.cuda.synchronize()
        for i in range(len(dev)):
            same_outputs(b.outputs(i), ref[i], f"tuned image {i}")
    finally:
        b.close()


def test_batch_rejects_alpha_and_mixed_fma(J, fe, torch):
    from jixel_amd import synth
    rgba = to_dev(torch, synth.smooth_rgba(40, 30, 1))
    p = J.frame_params(J.EncodeConfig())
    with pytest.raises(J.JxgError):
        J.Batch(fe, [(rgba, p)])
    rgb = to_dev(torch, synth.photo_like(40, 30, 1))
    q = J.frame_params(J.EncodeConfig(fused=False))
    with pytest.raises(J.JxgError):
        J.Batch(fe, [(rgb, p), (rgb, q)])
