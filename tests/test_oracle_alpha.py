"""Oracle of the alpha stages, pinned by the reference's own test vectors
(src/invisible.rs:158-237, src/alpha_fringe.rs:228-300,
src/adaptive_quant.rs:1076-1098) and by a direct-window brute force that
shares no bookkeeping with the reference's running-window sweep."""
import numpy as np
import pytest


def planes(w, h, values):
    v = np.asarray(values, np.float32).reshape(h, w)
    return np.stack([v, v, v])


# --- fill_invisible -------------------------------------------------------

def test_fill_propagates_in_scan_order(oracle):
    # src/invisible.rs:158-175: the filled upper-middle sample feeds the
    # lower-middle one, which feeds the last
    img = planes(3, 2, [2.0, 99.0, 4.0, 6.0, 99.0, 99.0])
    a = np.array([[255, 0, 1], [128, 0, 0]], np.uint8)
    out = oracle.fill_invisible(img, a)
    want = np.array([2.0, 3.0, 4.0, 6.0, 4.8, 4.266667], np.float32).reshape(2, 3)
    for c in range(3):
        np.testing.assert_allclose(out[c], want, atol=1e-6)
    # exact values: (2 * 2 + 4) / 3 ... written as the reference computes them
    assert out[0, 0, 1] == np.float32(9.0) / np.float32(3.0)


@pytest.mark.parametrize("w,h,alpha,src,want", [
    (1, 1, [0], [99.0], [0.0]),
    (3, 1, [0, 1, 0], [99.0, 2.0, 99.0], [2.0] * 3),
    (1, 3, [0, 1, 0], [99.0, 2.0, 99.0], [2.0] * 3),
    (2, 2, [0] * 4, [99.0] * 4, [0.0] * 4),
])
def test_fill_single_rows_columns_and_transparent(oracle, w, h, alpha, src, want):
    # src/invisible.rs:177-195
    out = oracle.fill_invisible(planes(w, h, src), np.asarray(alpha, np.uint8).reshape(h, w))
    for c in range(3):
        assert np.array_equal(out[c].reshape(-1), np.asarray(want, np.float32))


def test_fill_keeps_visible_and_constant_fields(oracle):
    rng = np.random.default_rng(5)
    h, w = 37, 53
    a = (rng.random((h, w)) < 0.6).astype(np.uint8) * rng.integers(1, 256, (h, w)).astype(np.uint8)
    a[0, 0] = 255  # every fill then has a neighbour
    img = rng.standard_normal((3, h, w)).astype(np.float32)
    out = oracle.fill_invisible(img, a)
    vis = a > 0
    assert np.array_equal(out[:, vis], img[:, vis])
    # a weighted mean of a constant (exact in binary) is that constant
    const = np.full((1, h, w), 0.5, np.float32)
    filled = oracle.fill_invisible(const, a)
    assert np.array_equal(filled, const)


def test_fill_without_invisible_is_identity(oracle):
    rng = np.random.default_rng(6)
    img = rng.standard_normal((3, 9, 11)).astype(np.float32)
    out = oracle.fill_invisible(img, np.full((9, 11), 7, np.uint8))
    assert np.array_equal(out, img)


# --- alpha_fringe::clear ----------------------------------------------------

def fringe_brute(a):
    """clear() restated as two direct 9x9 window stencils over the original
    samples, without the reference's running column sums or pending rows."""
    a = a.astype(np.int64)
    h, w = a.shape

    def sweep(src, inv, rep):
        lv = 255 - src if inv else src
        f = np.where(lv <= 4, lv, 0)
        out = src.copy()
        changed = 0
        for y, x in zip(*np.nonzero(f)):
            y0, y1 = max(0, y - 4), min(h, y + 5)
            x0, x1 = max(0, x - 4), min(w, x + 5)
            if f[y0:y1, x0:x1].sum() * 2 <= (y1 - y0) * (x1 - x0):
                out[y, x] = rep
                changed += 1
        return out, changed

    a1, _ = sweep(a, False, 0)
    d = 255 - a1
    if ((d >= 1) & (d <= 4)).sum() * 12 < a.size:
        return a1.astype(np.uint8)
    a2, changed = sweep(a1, True, 255)
    return (a2 if changed * 12 >= a.size else a1).astype(np.uint8)


def test_fringe_clears_thin_edges_keeps_shadow_tails(oracle):
    # src/alpha_fringe.rs:232-262
    w = h = 64
    y, x = np.mgrid[0:h, 0:w]
    a = np.zeros((h, w), np.uint8)
    a[:, :20] = 255
    a[:, 20] = 3
    a[20:51, 40:64] = 2
    out = oracle.alpha_fringe_clear(a)
    assert (out[:, 20] == 0).all()
    assert out[35, 52] == 2
    assert out[35, 10] == 255
    assert np.array_equal(out, fringe_brute(a))


def test_fringe_fills_broad_near_opaque_backdrop(oracle):
    # src/alpha_fringe.rs:264-299
    w = h = 64
    y, x = np.mgrid[0:h, 0:w]
    a = np.where(y <= 7, 128, np.where((x + 2 * y) % 4 == 0, 254, 255)).astype(np.uint8)
    out = oracle.alpha_fringe_clear(a)
    assert (out[8:] == 255).all() and (out[:8] == 128).all()
    sparse = np.full(w * h, 255, np.uint8)
    sparse[::40] = 254
    sparse = sparse.reshape(h, w)
    assert np.array_equal(oracle.alpha_fringe_clear(sparse), sparse)
    shade = np.full((h, w), 253, np.uint8)
    assert np.array_equal(oracle.alpha_fringe_clear(shade), shade)


@pytest.mark.parametrize("seed", range(8))
def test_fringe_matches_direct_window(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    h, w = int(rng.integers(1, 40)), int(rng.integers(1, 40))
    pick = rng.random((h, w))
    a = np.where(pick < 0.3, rng.integers(0, 6, (h, w)),
                 np.where(pick < 0.7, rng.integers(250, 256, (h, w)),
                          rng.integers(0, 256, (h, w)))).astype(np.uint8)
    if seed % 2:  # an opaque backdrop dithered by one level: the second sweep commits
        dither = np.where(rng.random((h, w)) < 0.3, 254, 255)
        a = np.where(rng.random((h, w)) < 0.9, dither, a).astype(np.uint8)
    out = oracle.alpha_fringe_clear(a)
    assert np.array_equal(out, fringe_brute(a))
    if seed % 2:
        assert ((a < 255) & (out == 255)).sum() > 0


# --- alpha block factors ----------------------------------------------------

def test_alpha_factors_follow_largest_alpha(oracle):
    # src/adaptive_quant.rs:1076-1098
    w, h = 24, 8
    a = np.zeros((h, w), np.uint8)
    a[:, 8:24] = 10
    a[3, 12] = 255
    f = oracle.alpha_block_factors(a)
    assert f.shape == (1, 3)
    assert f[0, 0] == 1.0 and f[0, 1] == 1.0
    assert abs(f[0, 2] - 0.1) < 1e-6
    a[3, 12] = 128
    f = oracle.alpha_block_factors(a)
    assert abs(f[0, 1] - 128.0 / 255.0) < 1e-6


def test_apply_alpha_factors_closed_form(oracle):
    from jixel_amd import STRATEGY as S
    rng = np.random.default_rng(9)
    xb, yb = 12, 8
    cells = np.full((yb, xb), (S["DCT"] << 1) | 1, np.uint8)
    # a DCT16X16 at (2, 2) and a DCT32X32 at (4, 4)
    for (bx, by, s, n) in [(2, 2, S["DCT16X16"], 2), (4, 4, S["DCT32X32"], 4)]:
        cells[by:by + n, bx:bx + n] = s << 1
        cells[by, bx] = (s << 1) | 1
    qf = rng.integers(1, 256, (yb, xb)).astype(np.uint8)
    f = rng.choice(np.float32([0.1, 0.35, 0.5, 0.77, 1.0]), (yb, xb)).astype(np.float32)
    out = oracle.apply_alpha_factors(qf, cells, f)
    want = qf.copy()
    for by in range(yb):
        for bx in range(xb):
            if not cells[by, bx] & 1:
                continue
            n = {S["DCT"]: 1, S["DCT16X16"]: 2, S["DCT32X32"]: 4}[cells[by, bx] >> 1]
            k = f[by:by + n, bx:bx + n].max()
            q = np.float32(qf[by, bx]) * np.float32(k)
            want[by:by + n, bx:bx + n] = min(max(int(np.floor(np.float64(q) + 0.5)), 1), 255)
    assert np.array_equal(out, want)


def test_front_end_alpha_plane_is_the_cleared_input(oracle):
    import jixel_amd as J
    from jixel_amd import synth
    rgba = synth.smooth_rgba(96, 80, 3)
    p = J.frame_params(J.EncodeConfig().with_quality(75))
    out = oracle.front_end(rgba, p, oracle.tables_for(p), threads=2)
    assert np.array_equal(out["alpha"], oracle.alpha_fringe_clear(rgba[:, :, 3]))
