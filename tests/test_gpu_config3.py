"""The 7680x4320 RGBA q75 workload with large transforms, at its full size.

* The whole front end (fringe clearing, fill_invisible, AQ, CfL maps,
  strategy, DCT + quantize, alpha factors) is bit-identical to the oracle's
  threaded front end on the same image.
* The DCT64X64 and DCT128X128 image stages, plain FP32 butterflies and
  split-bf16 x3 matrix cores, are checked block by block through properties
  that hold at any size, evaluated in FP64 on the device: the DC coefficient
  equals the block mean (the scaled DCT's normalisation), Parseval's identity
  sum X^2 = n^2 sum C^2 holds (C_n / sqrt n is orthogonal), and the split-bf16
  form stays within 3e-5 of the plain form relative to the block's largest
  coefficient (each is within 1e-5 / 2e-5 of FP64 on the smaller images of
  the large-transform tests).
"""
import numpy as np
import pytest

from test_gpu_parity import first_mismatch, same_bits, to_dev

pytestmark = pytest.mark.gpu

W, H, SEED, QUALITY = 7680, 4320, 3, 75.0


def log(*a):
    print(*a, flush=True)


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


@pytest.fixture(scope="module")
def rgba():
    from jixel_amd import synth
    log(f"synthesising the {W}x{H} RGBA image")
    img = synth.smooth_rgba(W, H, SEED)
    log("image ready")
    return img


def test_front_end_full_size_bit_exact(J, fe, torch, oracle, rgba):
    p = J.frame_params(J.EncodeConfig().with_quality(QUALITY))
    log(f"front end {W}x{H} RGBA on the GPU")
    o = fe.run(to_dev(torch, rgba), W, H, p)
    got = {k: o[k].cpu().numpy() for k in ("qf", "cells", "ytox", "ytob")}
    got_alpha = fe.alpha_plane(W, H).cpu().numpy()
    lv = [o["levels"][c].cpu().numpy() for c in range(3)]
    cf = [o["coeffs"][c].cpu().numpy() for c in range(3)]
    dc = [o["dc"][c].cpu().numpy() for c in range(3)]
    del o
    log("oracle front end (16 threads)")
    want = oracle.front_end(rgba, p, oracle.tables_for(p), threads=16)
    log("comparing")
    assert np.array_equal(got_alpha, want["alpha"])
    assert (want["alpha"] == 0).any() and (want["alpha"] == 255).any()
    for k, v in got.items():
        assert same_bits(v, want[k]), (k, first_mismatch(v, want[k]))
    for c in range(3):
        assert same_bits(lv[c], want["levels"][c]), ("levels", c, first_mismatch(lv[c], want["levels"][c]))
        assert same_bits(cf[c], want["coeffs"][c]), ("coeffs", c, first_mismatch(cf[c], want["coeffs"][c]))
        assert same_bits(dc[c], want["dc"][c]), ("dc", c, first_mismatch(dc[c], want["dc"][c]))


def blocks(torch, out, n, tiles_y, tiles_x):
    """(tiles_y, tiles_x, n*n) storage-order coefficients from the chunk
    layout (chunk j of a block = its 8-px row j of blocks, n/8 blocks x 64)."""
    k = n // 8
    o = out.view(tiles_y, k, tiles_x, k * 64)
    return o.permute(0, 2, 1, 3).reshape(tiles_y, tiles_x, n * n)


@pytest.mark.parametrize("n", [64, 128])
def test_large_transforms_full_size_properties(J, fe, torch, rgba, n):
    xyb = fe.xyb(to_dev(torch, rgba), W, H)
    assert xyb.shape[1] % n == 0 and xyb.shape[2] % n == 0
    ty, tx = (H + n - 1) // n, (W + n - 1) // n
    run = fe.dct64_image if n == 64 else fe.dct128_image
    log(f"DCT{n}X{n} of the {W}x{H} planes, plain and split-bf16")
    plain = run(xyb, W, H, 0)
    mfma = run(xyb, W, H, 1)
    mfma6 = run(xyb, W, H, 2) if n == 64 else None
    torch.cuda.synchronize()
    for c in range(3):
        x = xyb[c, :ty * n, :tx * n].double().view(ty, n, tx, n).permute(0, 2, 1, 3)
        mean = x.mean(dim=(2, 3))
        energy = (x * x).sum(dim=(2, 3))
        p = blocks(torch, plain[c], n, ty, tx).double()
        m = blocks(torch, mfma[c], n, ty, tx).double()
        scale = p.abs().amax(dim=2)
        # saturated blocks (R = G = B) have X = 0 exactly: all-zero coefficients
        zero = scale == 0
        log(f"  c{c}: {int(zero.sum())} all-zero blocks of {ty * tx}")
        assert bool((energy[zero] == 0).all()) and bool((m.abs().amax(dim=2)[zero] == 0).all())
        scale = torch.where(zero, torch.ones_like(scale), scale)
        denom = torch.where(energy == 0, torch.ones_like(energy), energy)
        for name, y in (("plain", p), ("split-bf16", m)):
            dc_err = ((y[:, :, 0] - mean).abs() / scale).max().item()
            pars = ((n * n * (y * y).sum(dim=2) - energy).abs() / denom).max().item()
            log(f"  c{c} {name}: DC vs mean {dc_err:.2e}, Parseval {pars:.2e}")
            assert dc_err <= (1e-5 if name == "plain" else 2e-5)
            assert pars <= (1e-5 if name == "plain" else 4e-5)
        diff = ((m - p).abs().amax(dim=2) / scale).max().item()
        log(f"  c{c} split-bf16 vs plain {diff:.2e}")
        assert diff <= 3e-5
        if n == 64:
            # the three-term split with six products
            m6 = blocks(torch, mfma6[c], n, ty, tx).double()
            d6 = ((m6 - p).abs().amax(dim=2) / scale).max().item()
            pars6 = ((n * n * (m6 * m6).sum(dim=2) - energy).abs() / denom).max().item()
            log(f"  c{c} split-bf16x6 vs plain {d6:.2e}, Parseval {pars6:.2e}")
            assert d6 <= 1e-5 and pars6 <= 1e-5
