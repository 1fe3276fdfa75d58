"""Oracle pins for the EPF luma proxy (LumaProxy, src/epf_proxy.rs:151-245).

The box moments are checked against a direct FP64 5x5 mirrored mean, the
gradient pools against an FP64 vectorised form, identity and flat-field
behaviour against the reference's own test (src/epf_proxy.rs:266-283), and
the f32 pixel-order pooling against the f64 pooling the GPU form uses."""
import numpy as np
import pytest


def mirror(i, n):
    if 0 <= i < n:
        return i
    m = i % (2 * n)
    return m if m < n else 2 * n - m - 1


def box5_f64(src):
    h, w = src.shape
    ry = np.array([[mirror(y + d, h) for d in range(-2, 3)] for y in range(h)])
    rx = np.array([[mirror(x + d, w) for d in range(-2, 3)] for x in range(w)])
    s = src.astype(np.float64)
    v = s[ry].sum(axis=1)          # (h, w): vertical 5-sums
    return v[:, rx].sum(axis=2) / 25.0


@pytest.mark.parametrize("w,h", [(1, 1), (3, 2), (7, 5), (33, 17), (130, 66)])
def test_box5_matches_fp64(oracle, w, h):
    rng = np.random.default_rng(w * 31 + h)
    src = rng.uniform(-0.3, 0.9, (h, w)).astype(np.float32)
    got = oracle.epf_box5(src).astype(np.float64)
    assert np.abs(got - box5_f64(src)).max() < 2e-6


def grad_terms_f64(src, y):
    s, c = src.astype(np.float64), y.astype(np.float64)
    out = []
    for axis in (1, 0):
        a = np.diff(s, axis=axis)
        b = np.diff(c, axis=axis)
        out.append((np.abs(b - a).sum(), np.maximum(np.abs(b) - np.abs(a), 0).sum(),
                    (np.maximum(np.abs(a) - np.abs(b), 0) ** 4).sum()))
    h, w = src.shape
    n = h * (w - 1) + (h - 1) * w
    e, ex, l4 = [(out[0][i] + out[1][i]) for i in range(3)]
    return e / n, ex / n, (l4 / n) ** 0.25


def test_identity_scores_zero_and_shift_scores_higher(oracle):
    """src/epf_proxy.rs:266-283 on flat fields, plus textured content."""
    for value in [0.0, 0.07, 0.21, 0.53, 0.91]:
        y = np.full((129, 257), value, np.float32)
        ident, _ = oracle.epf_compare(y, y)
        assert ident < 1e-5
        shifted, _ = oracle.epf_compare(y, y + np.float32(0.02))
        assert np.isfinite(shifted) and shifted > ident
    rng = np.random.default_rng(5)
    y = rng.uniform(0, 1, (96, 128)).astype(np.float32)
    ident, terms = oracle.epf_compare(y, y)
    assert ident == 0.0 and not terms.any()


def test_terms_against_fp64(oracle):
    rng = np.random.default_rng(6)
    src = rng.uniform(0.1, 0.8, (72, 100)).astype(np.float32)
    cand = (src + rng.normal(0, 0.02, src.shape)).astype(np.float32)
    score, terms = oracle.epf_compare(src, cand, wide=True)
    e, ex, l4 = grad_terms_f64(src, cand)
    assert np.allclose(terms[3:], [e, ex, l4], rtol=1e-5)
    assert 0 < terms[0] < 1 and 0 < terms[1] < 1 and 0 < terms[2] < 1
    w = np.array([109.4163228369586, 4.685925102200196, 2.491781270704224, 2.167275675983448,
                  4.55227651921194, 2.847714678987977])
    assert abs(score - float(terms.astype(np.float64) @ w)) < 1e-5 * score


def test_f32_and_f64_pooling_agree(oracle):
    rng = np.random.default_rng(7)
    src = rng.uniform(0.1, 0.8, (256, 256)).astype(np.float32)
    cand = (src + rng.normal(0, 0.01, src.shape)).astype(np.float32)
    s32, t32 = oracle.epf_compare(src, cand, wide=False)
    s64, t64 = oracle.epf_compare(src, cand, wide=True)
    assert np.allclose(t32, t64, rtol=1e-4) and abs(s32 - s64) < 1e-4 * s64
