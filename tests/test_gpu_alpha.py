"""GPU parity of the alpha stages (fringe clearing, fill_invisible, alpha
block factors and their application) and of the RGBA / gray + alpha front
end, against the oracle, bit for bit."""
import numpy as np
import pytest

from test_gpu_parity import first_mismatch, same_bits, to_dev

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


def fringe_plane(h, w, seed):
    """Alpha with fringe samples, shadow tails and dithered near-opaque areas."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    a = np.where((x // 7 + y // 5) % 3 == 0, 255, 0)
    a = np.where(rng.random((h, w)) < 0.15, rng.integers(1, 5, (h, w)), a)
    a = np.where((y > h // 2) & (rng.random((h, w)) < 0.8),
                 np.where(rng.random((h, w)) < 0.3, 254, 255), a)
    return a.astype(np.uint8)


def ellipse_alpha(h, w, seed, zero_share=0.5):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    a = np.zeros((h, w))
    for _ in range(6):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        rx, ry = rng.uniform(3, w / 2 + 3), rng.uniform(3, h / 2 + 3)
        d = np.sqrt(((x - cx) / rx) ** 2 + ((y - cy) / ry) ** 2)
        a = np.maximum(a, np.clip((1.0 - d) * 8.0, 0.0, 1.0))
    a = np.round(a * 255).astype(np.uint8)
    return np.where(rng.random((h, w)) < zero_share * 0.02, 0, a).astype(np.uint8)


@pytest.mark.parametrize("w,h,seed", [(64, 64, 1), (67, 45, 2), (200, 130, 3), (5, 300, 4),
                                      (333, 7, 5), (1, 1, 6)])
def test_fringe_clear_bit_exact(fe, torch, oracle, w, h, seed):
    log("fringe", w, h, seed)
    a = fringe_plane(h, w, seed)
    want = oracle.alpha_fringe_clear(a)
    d = to_dev(torch, a)
    fe.alpha_fringe_clear(d, w, h)
    got = d.cpu().numpy()
    assert np.array_equal(got, want), first_mismatch(got, want)


def test_fringe_clear_reference_cases(fe, torch, oracle):
    # the two pictures of src/alpha_fringe.rs's tests: a cleared thin edge,
    # and a dithered backdrop that the second sweep fills
    y, x = np.mgrid[0:64, 0:64]
    a = np.zeros((64, 64), np.uint8)
    a[:, :20] = 255
    a[:, 20] = 3
    a[20:51, 40:64] = 2
    b = np.where(y <= 7, 128, np.where((x + 2 * y) % 4 == 0, 254, 255)).astype(np.uint8)
    for plane in (a, b):
        d = to_dev(torch, plane)
        fe.alpha_fringe_clear(d, 64, 64)
        assert np.array_equal(d.cpu().numpy(), oracle.alpha_fringe_clear(plane))


def _xyb_planes(J, fe, torch, w, h, seed):
    rng = np.random.default_rng(seed)
    rgb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    xyb = fe.xyb(to_dev(torch, rgb), w, h)
    return xyb


@pytest.mark.parametrize("w,h,seed,kind", [
    (37, 29, 1, "ellipses"), (130, 70, 2, "random"), (1, 1, 3, "zero"), (1, 9, 4, "random"),
    (9, 1, 5, "random"), (64, 64, 6, "zero"), (200, 1100, 7, "ellipses"),
    (41, 2100, 8, "random"), (300, 40, 9, "opaque"), (257, 129, 10, "sparse")])
def test_fill_invisible_bit_exact(J, fe, torch, oracle, w, h, seed, kind):
    log("fill_invisible", w, h, seed, kind)
    rng = np.random.default_rng(seed)
    if kind == "ellipses":
        a = ellipse_alpha(h, w, seed)
    elif kind == "zero":
        a = np.zeros((h, w), np.uint8)
    elif kind == "opaque":
        a = np.full((h, w), 255, np.uint8)
    elif kind == "sparse":
        a = np.where(rng.random((h, w)) < 0.05, 0, 200).astype(np.uint8)
    else:
        a = np.where(rng.random((h, w)) < 0.5, 0, rng.integers(1, 256, (h, w))).astype(np.uint8)
    for y in range(1024, h, 1024):  # a pass's first row reads the sample above it
        a[y, 0], a[y - 1, 0] = 0, 255
    xyb = _xyb_planes(J, fe, torch, w, h, seed)
    src = xyb[:, :h, :w].cpu().numpy()
    want = oracle.fill_invisible(src, a)
    fe.fill_invisible(xyb, w, h, to_dev(torch, a))
    got = xyb.cpu().numpy()
    for c in range(3):
        assert same_bits(got[c, :h, :w], want[c]), (c, first_mismatch(got[c, :h, :w], want[c]))
        # edge padding replicates the filled edge samples
        s, ph = J.plane_geometry(w, h)
        pad = np.pad(want[c], ((0, ph - h), (0, s - w)), mode="edge")
        assert same_bits(got[c], pad), c


@pytest.mark.parametrize("w,h,seed,scale", [(70, 90, 11, 2.0 ** -118), (150, 200, 12, 2.0 ** -126),
                                             (33, 140, 13, 2.0 ** -140)])
def test_fill_invisible_tiny_samples_bit_exact(J, fe, torch, oracle, w, h, seed, scale):
    """Samples near and below the smallest normal float: some weighted sums
    fall where the reciprocal division is not exact, and the fill must still
    equal IEEE division (the oracle's) bit for bit.  The first two rows of
    bands keep ordinary values, so only some bands meet the tiny sums."""
    log("fill_invisible tiny", w, h, seed, scale)
    rng = np.random.default_rng(seed)
    a = np.where(rng.random((h, w)) < 0.5, 0, 255).astype(np.uint8)
    s, ph = J.plane_geometry(w, h)
    planes = rng.uniform(-1.0, 1.0, (3, ph, s))
    planes[:, min(h, 128):, :] *= scale
    xyb = torch.from_numpy(planes.astype(np.float32)).cuda()
    src = xyb[:, :h, :w].cpu().numpy()
    assert np.count_nonzero(np.abs(src) < 2.0 ** -124) > 0 or scale > 2.0 ** -120
    want = oracle.fill_invisible(src, a)
    fe.fill_invisible(xyb, w, h, to_dev(torch, a))
    got = xyb.cpu().numpy()
    for c in range(3):
        assert same_bits(got[c, :h, :w], want[c]), (c, first_mismatch(got[c, :h, :w], want[c]))


@pytest.mark.parametrize("offset", [1, 3, 13])
def test_fill_invisible_unaligned_alpha(J, fe, torch, oracle, offset):
    """An alpha plane that does not start on a 16-byte boundary (the stats
    pass reads 16-byte words from the first aligned address)."""
    w, h = 173, 91
    log("fill_invisible unaligned alpha", offset)
    rng = np.random.default_rng(offset)
    a = np.where(rng.random((h, w)) < 0.3, 0, 255).astype(np.uint8)
    buf = torch.zeros(w * h + 32, dtype=torch.uint8, device="cuda")
    buf[offset:offset + w * h] = torch.from_numpy(a.reshape(-1)).cuda()
    xyb = _xyb_planes(J, fe, torch, w, h, offset)
    src = xyb[:, :h, :w].cpu().numpy()
    want = oracle.fill_invisible(src, a)
    fe.fill_invisible(xyb, w, h, buf[offset:offset + w * h])
    got = xyb.cpu().numpy()
    for c in range(3):
        assert same_bits(got[c, :h, :w], want[c]), (c, first_mismatch(got[c, :h, :w], want[c]))


@pytest.mark.parametrize("w,h,seed", [(24, 8, 1), (131, 77, 2), (64, 64, 3)])
def test_alpha_factors_bit_exact(J, fe, torch, oracle, w, h, seed):
    rng = np.random.default_rng(seed)
    a = ellipse_alpha(h, w, seed)
    a = np.where(rng.random((h, w)) < 0.1, rng.integers(0, 256, (h, w)), a).astype(np.uint8)
    want = oracle.alpha_block_factors(a)
    got = fe.alpha_block_factors(to_dev(torch, a), w, h).cpu().numpy()
    assert same_bits(got, want), first_mismatch(got, want)
    # apply on a layout with every square size
    from jixel_amd import STRATEGY as S
    xb, yb = (w + 7) // 8, (h + 7) // 8
    cells = np.full((yb, xb), (S["DCT"] << 1) | 1, np.uint8)
    for by in range(0, yb - 3, 4):
        for bx in range(0, xb - 3, 8):
            n = 4 if (bx // 8 + by // 4) % 2 else 2
            cells[by:by + n, bx:bx + n] = S["DCT32X32" if n == 4 else "DCT16X16"] << 1
            cells[by, bx] |= 1
    qf = rng.integers(1, 256, (yb, xb)).astype(np.uint8)
    want_q = oracle.apply_alpha_factors(qf, cells, want)
    d_q = to_dev(torch, qf)
    fe.apply_alpha_factors(to_dev(torch, cells), to_dev(torch, want), w, h, d_q)
    assert np.array_equal(d_q.cpu().numpy(), want_q)


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h,ch,q,flags", [
    (160, 96, 4, 75.0, 3), (203, 117, 4, 90.0, 3), (136, 72, 2, 80.0, 3),
    (120, 64, 4, 75.0, 1), (120, 64, 4, 75.0, 0), (96, 1030, 4, 60.0, 3)])
def test_front_end_alpha_bit_exact(J, fe, torch, oracle, w, h, ch, q, flags, fused):
    log("front_end alpha", w, h, ch, q, flags, fused)
    from jixel_amd import synth
    rgba = synth.smooth_rgba(w, h, 11)
    a = rgba[:, :, 3].copy()
    rng = np.random.default_rng(w + h)
    a = np.where(rng.random((h, w)) < 0.05, rng.integers(1, 5, (h, w)), a).astype(np.uint8)
    img = rgba.copy()
    img[:, :, 3] = a
    if ch == 2:
        img = np.ascontiguousarray(np.stack([img[:, :, 1], a], axis=2))
    p = J.frame_params(J.EncodeConfig(fused=fused).with_quality(q))
    p.flags = flags
    want = oracle.front_end(img, p, oracle.tables_for(p), threads=4)
    o = fe.run(to_dev(torch, img), w, h, p)
    got_alpha = fe.alpha_plane(w, h).cpu().numpy()
    assert np.array_equal(got_alpha, want["alpha"])
    for k in ("qf", "cells", "ytox", "ytob"):
        assert same_bits(o[k].cpu().numpy(), want[k]), (k, first_mismatch(o[k].cpu().numpy(), want[k]))
    for c in range(3):
        assert same_bits(o["levels"][c].cpu().numpy(), want["levels"][c]), c
        assert same_bits(o["coeffs"][c].cpu().numpy(), want["coeffs"][c]), c
        assert same_bits(o["dc"][c].cpu().numpy(), want["dc"][c]), c


def test_host_front_end_returns_alpha(J, oracle):
    from jixel_amd import synth
    rgba = synth.smooth_rgba(90, 70, 12)
    cfg = J.EncodeConfig().with_quality(75)
    out = J.encode_front_end(rgba, cfg)
    want = oracle.front_end(rgba, J.frame_params(cfg), threads=2)
    assert np.array_equal(out.alpha, want["alpha"])
    assert same_bits(out.qf, want["qf"]) and same_bits(out.cells, want["cells"])
