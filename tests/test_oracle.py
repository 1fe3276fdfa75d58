"""Pins the CPU oracle against the reference's own tables, known answers and
closed forms (CPU only)."""
import ctypes
import math
import struct

import numpy as np
import pytest


def f32(x):
    return float(np.float32(x))


def bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


# ---- tables the reference stores --------------------------------------------

@pytest.mark.parametrize("name,cx,cy", [
    ("K_COEFF_ORDER_8X8", 1, 1), ("K_COEFF_ORDER_16X8", 2, 1), ("K_COEFF_ORDER_16X16", 2, 2),
    ("K_COEFF_ORDER_32X16", 4, 2), ("K_COEFF_ORDER_32X32", 4, 4)])
def test_natural_scan_matches_reference_orders(oracle, golden, name, cx, cy):
    # src/ac_context.rs:164-340 vs natural_scan_generic (src/coeff_order.rs:206)
    assert np.array_equal(oracle.natural_scan(cx, cy), np.array(golden[name], np.uint32))


def test_scan_pos_lut_is_inverse_and_transposes(oracle):
    for w, h in [(8, 8), (16, 8), (8, 16), (32, 32), (32, 16), (16, 32), (32, 8), (8, 32)]:
        lut = oracle.scan_pos_lut(w, h)
        assert sorted(lut.tolist()) == list(range(w * h))
        if w < h:  # tall = transpose of the wide table (src/coeff_order.rs:395-404)
            wide = oracle.scan_pos_lut(h, w).reshape(w, h)
            assert np.array_equal(lut.reshape(h, w), wide.T)


def test_exp2_table_matches_reference(oracle, golden):
    import ctypes
    L = oracle.lib()
    L.jxo_exp2_table.restype = ctypes.POINTER(ctypes.c_float)
    t = L.jxo_exp2_table()
    assert [bits(t[i]) for i in range(64)] == [p[0] for p in golden["EXP2FT"]]


def test_butterfly_constants_are_the_reference_constants(oracle, golden):
    # WC_N[i] = 1 / (2 cos((2i+1) pi / 2N)) to within an f32 ulp (src/dct.rs:55-59,2404)
    for n, name in [(8, "WC8"), (16, "WC16"), (32, "WC32"), (64, "WC64")]:
        for i, v in enumerate(golden[name]):
            exact = 1.0 / (2.0 * math.cos((2 * i + 1) * math.pi / (2 * n)))
            assert abs(f32(v) - exact) <= 2e-7 * exact, (name, i)
    L = oracle.lib()
    L.jxo_wc64.restype = ctypes.POINTER(ctypes.c_float)
    w = L.jxo_wc64()
    assert [bits(w[i]) for i in range(32)] == [bits(f32(v)) for v in golden["WC64"]]


# ---- known answers --------------------------------------------------------

def test_quality_mapping_anchor_points(oracle):
    # src/encode_image.rs:5758-5771
    d = lambda q: oracle.scalar("distance_from_quality", q)
    close = lambda a, b: abs(a - b) < 1e-4
    assert close(d(100.0), 0.05) and close(d(99.0), 0.1) and close(d(90.0), 1.0)
    assert close(d(30.0), 6.4)
    assert 6.4 < d(25.0) < 7.0
    assert close(d(0.0), 25.0) and close(d(-50.0), 25.0) and close(d(110.0), 0.05)
    prev = d(0.0)
    for q in range(1, 101):
        cur = d(float(q))
        assert cur <= prev
        prev = cur


def test_distance_params_at_q75(oracle):
    # global_scale = clamp(round(65536*0.8/(d*10)), 1, 2^15) = 2231 at d=2.35
    p = oracle.distance_params(f32(2.35))
    assert p["global_scale"] == 2231
    assert p["scale"] == 2231 / 65536.0
    assert p["x_qm_scale"] == 3 and p["epf_iters"] == 2
    p1 = oracle.distance_params(1.0)
    assert p1["global_scale"] == round(65536 * 0.8 / 10.0) and p1["x_qm_scale"] == 2


def test_dirty_log2f_exact_at_powers_of_two_and_bounded(oracle):
    # src/adaptive_quant.rs:896-917
    for e in range(-20, 21):
        assert oracle.scalar("dirty_log2f", 2.0 ** e) == float(e)
    x = 1e-6
    worst = 0.0
    while x <= 1e6:
        worst = max(worst, abs(oracle.scalar("dirty_log2f", x) - math.log2(f32(x))))
        x *= 1.01
    assert worst < 2e-6


def test_fast_exp_within_two_ulp(oracle):
    # src/util.rs:360 (<= 2 ulp over its domain)
    for x in np.linspace(-80.0, 80.0, 4001):
        x = f32(x)
        got = oracle.scalar("fast_exp", x)
        want = f32(math.exp(x))
        ulp = abs(bits(got) - bits(want))
        assert ulp <= 2, (x, got, want)


def test_fast_exp2_close_to_exp2(oracle):
    for x in np.linspace(-20.0, 20.0, 2001):
        x = f32(x)
        got = oracle.scalar("fast_exp2", x)
        assert abs(got - 2.0 ** x) <= 4e-7 * 2.0 ** x, x


def test_dequant_levels_match_the_division(oracle):
    # src/group.rs:2706-2731: level(q) == q - numerator/q for |q| >= 2
    L = oracle.lib()
    one, num = f32(1.0 - 0.07005449891748593), f32(0.145)
    for q in range(-300, 301):
        got = L.jxo_dequant_level(one, num, q)
        if q == 0:
            want = 0.0
        elif abs(q) == 1:
            want = math.copysign(one, q)
        else:
            want = f32(np.float32(q) - np.float32(num) / np.float32(q))
        assert bits(got) == bits(f32(want)), q


# ---- closed forms ------------------------------------------------------------

def _dct_mat(n):
    k = np.arange(n)[:, None]
    i = np.arange(n)[None, :]
    m = np.cos(np.pi * k * (2 * i + 1) / (2 * n))
    m[1:] *= math.sqrt(2.0)
    return m


SHAPES = {0: (8, 8), 1: (8, 16), 2: (16, 8), 3: (16, 16), 4: (32, 32), 8: (16, 32),
          9: (32, 16), 22: (8, 32), 23: (32, 8), 14: (64, 64), 15: (32, 64), 16: (64, 32)}


@pytest.mark.parametrize("strategy", sorted(SHAPES))
def test_dct_matches_fp64_closed_form(oracle, strategy):
    """X[v,u] = c_v c_u / (W H) sum x cos cos, in the reference's storage
    (square 8x8 u-major, wide v-major, column-first u-major)."""
    w, h = SHAPES[strategy]
    rng = np.random.default_rng(strategy)
    img = rng.uniform(-1, 1, (h, w)).astype(np.float32)
    out = oracle.forward_transform(strategy, img, 0, 0).astype(np.float64)
    ref = _dct_mat(h) @ img.astype(np.float64) @ _dct_mat(w).T / (w * h)
    exp = ref.reshape(-1) if w > h else ref.T.reshape(-1)
    assert np.abs(out - exp).max() <= 1e-5 * np.abs(exp).max()


def test_dct_edge_replication(oracle):
    # a footprint past the image repeats the last column / row
    rng = np.random.default_rng(3)
    plane = rng.uniform(0, 1, (20, 13)).astype(np.float32)
    got = oracle.forward_transform(3, plane, 8, 8)
    pad = np.pad(plane, ((0, 12), (0, 19)), mode="edge")
    want = oracle.forward_transform(3, np.ascontiguousarray(pad[8:24, 8:24]), 0, 0)
    assert np.array_equal(got, want)


def test_dc_of_constant_block_is_the_value(oracle):
    # every DC extraction of a flat footprint returns the pixel value
    L = oracle.lib()
    import ctypes
    for s, (w, h) in SHAPES.items():
        img = np.full((h, w), 0.375, np.float32)
        co = oracle.forward_transform(s, img, 0, 0)
        dc = np.zeros(64, np.float32)
        L.jxo_dc_from_coefficients(s, co.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   dc.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        n = (w // 8) * (h // 8)
        assert np.allclose(dc[:n], 0.375, atol=4e-6), (s, dc[:n])


def test_xyb_closed_form(oracle):
    """XYB against exact f64 cube roots (SURVEY closed form): <= 1e-6 rel."""
    from jixel_amd import XYB_SPEC_FWD
    lut = oracle.srgb_lut()
    m = XYB_SPEC_FWD.astype(np.float64).reshape(3, 3)
    rng = np.random.default_rng(0)
    rgb = rng.integers(0, 256, size=(1, 4096, 3), dtype=np.uint8)
    got = oracle.xyb_image(rgb, XYB_SPEC_FWD)
    lin = lut[rgb[0]].astype(np.float64)
    mixed = np.maximum(lin @ m.T + 0.0037930734, 0.0)
    tm = np.cbrt(mixed) - 0.1559542
    want = np.stack([0.5 * (tm[:, 0] - tm[:, 1]), 0.5 * (tm[:, 0] + tm[:, 1]), tm[:, 2]])
    assert np.abs(got[:, 0, :] - want).max() < 3e-6


def test_srgb_lut_closed_form(oracle):
    lut = oracle.srgb_lut()
    v = np.arange(256) / 255.0
    want = np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4).astype(np.float32)
    assert np.array_equal(lut, want)


def test_flat_quant_field_above_1_75(oracle):
    """For d >= 1.75 the masking strength is 0, so every block gets
    (0.48 * (0.8294/d) * inv_scale + 0.5) truncated (src/adaptive_quant.rs:693-726)."""
    from jixel_amd.synth import photo_like
    rgb = photo_like(96, 72, 7)
    from jixel_amd import XYB_SPEC_FWD
    xyb = oracle.xyb_image(rgb, XYB_SPEC_FWD)
    for q in (75.0, 60.0):
        d = oracle.scalar("distance_from_quality", q)
        p = oracle.distance_params(d)
        qf = oracle.fill_quant_field(xyb, d, 1.0 / p["scale"])
        scale = np.float32(0.8294) / np.float32(d)
        base = np.float32(np.float32(0.48) * scale)
        want = int(np.float32(np.float32(base * np.float32(1.0 / p["scale"])) + np.float32(0.5)))
        assert (qf == want).all(), (q, np.unique(qf), want)
    assert want == 5 or q != 75.0


def test_quantizer_brute_force(oracle):
    """quantize_block_ac is the textbook dead-zone quantizer with the
    quadrant thresholds (brute force over tie and threshold boundaries)."""
    import ctypes
    L = oracle.lib()
    for c in range(3):
        for (xs, ys) in [(1, 1), (2, 1), (2, 2), (4, 4)]:
            thr = np.zeros(4, np.float32)
            L.jxo_ac_thresholds(c, xs, ys, f32(2.35), f32(1.25 if c == 0 else 1.0),
                                thr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            n = xs * ys * 64
            w = xs * 8
            rng = np.random.default_rng(c * 10 + xs)
            vals = rng.choice(np.float32([0.0, 0.49999997, 0.5, 1.5, 2.5, -0.5, -2.5,
                                          *thr, *(-thr), *(thr - 1e-7)]), n).astype(np.float32)
            qm = np.ones(n, np.float32)
            out = np.zeros(n, np.int32)
            fp = ctypes.POINTER(ctypes.c_float)
            L.jxo_quantize_block_ac.argtypes = [fp, ctypes.c_int, fp, ctypes.c_int32, ctypes.c_float,
                                                ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                                ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
            L.jxo_quantize_block_ac(vals.ctypes.data_as(fp), c, qm.ctypes.data_as(fp), 1, 1.0, 1.0,
                                    f32(2.35), xs, ys,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
            for i in range(n):
                y, x = divmod(i, w)
                t = thr[(2 if y >= ys * 4 else 0) + (1 if x >= w // 2 else 0)]
                v = vals[i]
                want = 0 if abs(v) < t else int(math.copysign(math.floor(abs(v) + 0.5), v))
                assert out[i] == want, (c, xs, i, v, t)


@pytest.mark.parametrize("strategy", [14, 15, 16])
def test_dct64_dc_is_the_block_mean(oracle, strategy):
    """DC of each covered 8x8 block equals that block's mean for content
    whose spectrum lives in the lowest 8x8 frequencies (a smooth ramp plus
    one cosine of the 8-block grid period), which pins the resample scales,
    the inverse DCT8 and the output order (src/dct.rs:2602,2789,2800)."""
    w, h = SHAPES[strategy]
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = (0.2 + 0.3 * np.cos(np.pi * (2 * xx + 1) * 2 / (2 * w))
           + 0.25 * np.cos(np.pi * (2 * yy + 1) * 1 / (2 * h))).astype(np.float32)
    co = oracle.forward_transform(strategy, img, 0, 0)
    L = oracle.lib()
    dc = np.zeros(64, np.float32)
    L.jxo_dc_from_coefficients(strategy, co.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               dc.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    bw, bh = w // 8, h // 8
    means = img.astype(np.float64).reshape(bh, 8, bw, 8).mean(axis=(1, 3)).reshape(-1)
    assert np.abs(dc[:bw * bh] - means).max() < 2e-3, (dc[:bw * bh], means)


def test_gray_input_is_snapped_achromatic(oracle):
    """Gray samples decode to R = G = B (linearize_gray) and snap to X = 0,
    B = Y (snap_achromatic_xyb, src/frame/encoder.rs:347-357); Y equals the
    Y of the same gray as RGB."""
    import jixel_amd as J
    rng = np.random.default_rng(9)
    g = rng.integers(0, 256, (5, 7, 1), dtype=np.uint8)
    ga = np.concatenate([g, rng.integers(0, 256, (5, 7, 1), dtype=np.uint8)], axis=2)
    rgb = np.repeat(g, 3, axis=2)
    x1 = oracle.xyb_image(g, J.XYB_SPEC_FWD)
    x2 = oracle.xyb_image(ga, J.XYB_SPEC_FWD)
    x3 = oracle.xyb_image(rgb, J.XYB_SPEC_FWD)
    assert np.array_equal(x1, x2)
    assert not x1[0].any() and np.array_equal(x1[2], x1[1])
    assert np.array_equal(x1[1], x3[1])
    assert np.abs(x3[0]).max() < 1e-6 and np.abs(x3[2] - x3[1]).max() < 1e-6
