"""Oracle of the chroma-HF protection (src/adaptive_quant.rs:222-408), pinned
by closed-form block statistics, the fit's fade schedule and a float64
statistics restatement on the reference's own test picture (:812-853)."""
import numpy as np
import pytest


def stats64(xyb, px, py, w, h):
    x, y, b = (xyb[c, py:py + h, px:px + w].astype(np.float64) for c in range(3))
    o = b - y
    s = np.abs(x).sum() + np.abs(o).sum()
    g = np.abs(np.diff(x, axis=1)).sum() + np.abs(np.diff(o, axis=1)).sum()
    g += np.abs(np.diff(x, axis=0)).sum() + np.abs(np.diff(o, axis=0)).sum()
    lg = np.abs(np.diff(y, axis=1)).sum() + np.abs(np.diff(y, axis=0)).sum()
    return s, g, lg, (w - 1) * h + w * (h - 1)


def test_stats_closed_form(oracle):
    # X: vertical stripes +-a; Y: rows rising by g; B = Y + b0
    a, g, b0 = 0.0625, 0.03125, 0.125
    yy, xx = np.mgrid[0:8, 0:8]
    X = np.where(xx % 2, a, -a)
    Y = g * yy
    xyb = np.stack([X, Y, Y + b0]).astype(np.float32)
    s, cg, lg, e = oracle.chroma_hf_stats(xyb, 0, 0, 8, 8)
    assert s == 64 * (a + b0)
    assert cg == 56 * 2 * a       # horizontal X steps only
    assert lg == 56 * g           # vertical Y steps only
    assert e == 112


def test_stats_match_float64_on_reference_picture(oracle):
    # the picture of selected_chroma_hf_stats_match_scalar
    h, w = 17, 19
    xyb = np.zeros((3, h, w), np.float32)
    for c in range(3):
        for y in range(h):
            for x in range(w):
                bits = (x * 37 + y * 53 + c * 97 + x * y * 3) % 257
                xyb[c, y, x] = (np.float32(bits) - np.float32(128.0)) / np.float32(211.0)
    for py in (0, 3, 9, 16):
        for px in (0, 2, 7, 14, 18):
            for hh in range(1, min(h - py, 8) + 1):
                for ww in range(1, min(w - px, 8) + 1):
                    got = oracle.chroma_hf_stats(xyb, px, py, ww, hh)
                    want = stats64(xyb, px, py, ww, hh)
                    assert got[3] == want[3]
                    for k in range(3):
                        assert abs(got[k] - want[k]) <= 2e-5 * max(abs(want[k]), 1.0)


@pytest.mark.parametrize("d", [0.1, 0.35, 0.5, 0.75, 0.9, 1.0, 1.3, 1.99, 2.0, 2.35, 4.0])
def test_strength_schedule(oracle, d):
    fade_in = min(max((d - 0.35) / 0.40, 0.0), 1.0)
    fade_out = 1.0 - min(max((d - 1.0) / 1.0, 0.0), 1.0)
    want = 0.40 * fade_in * fade_out if d >= 0.35 else 0.0
    assert abs(oracle.chroma_hf_strength(d) - want) < 1e-6


@pytest.mark.parametrize("fused", [True, False])
def test_protection_closed_form(oracle, fused):
    # one strongly chromatic striped block, one flat block, one luma-only block
    a, g = 0.0625, 0.03125
    yy, xx = np.mgrid[0:8, 0:24]
    X = np.where((xx < 8) & (xx % 2 == 1), a, np.where(xx < 8, -a, 0.0))
    Y = np.where(xx >= 16, g * (xx % 2), 0.0)
    B = Y + np.where(xx < 8, 0.125, 0.0)
    xyb = np.stack([X, Y, B]).astype(np.float32)
    qf = np.array([[40, 40, 40]], np.uint8)
    d = 0.8
    oracle.set_fused(fused)
    out = oracle.chroma_hf_protection(xyb, qf, d)
    oracle.set_fused(True)
    ds = 0.40 * min((d - 0.35) / 0.40, 1.0) * 1.0
    mean_chroma = a + 0.125
    mean_grad = 112 * a / 112
    share = 1.0 / (1.0 + 1e-4 / (112 * a))
    cw = min(max((mean_chroma - 0.025) / 0.075, 0), 1)
    gw = min(max((mean_grad - 0.006) / 0.030, 0), 1)
    sw = min(max((share - 0.25) / 0.55, 0), 1)
    gain = 1.0 + (0.35 + 1.4 * sw * sw) * ds * cw * gw * sw
    assert out[0, 0] == int(np.floor(40 * gain + 0.5)) and out[0, 0] > 40
    assert out[0, 1] == 40 and out[0, 2] == 40
    # outside the fade band nothing changes
    assert np.array_equal(oracle.chroma_hf_protection(xyb, qf, 2.35), qf)
    assert np.array_equal(oracle.chroma_hf_protection(xyb, qf, 0.3), qf)
