"""The default quant tables: DequantMatrices::new_fast(distance, arbiter)
(src/quant_weights.rs:2294-2337), restated twice (the product's
jxg_default_quant_tables in jixel_amd/csrc/jxg_tables.cpp and the oracle's
jxo_fast_tables in oracle/jxo_tables.c, which rounds f16 through the F16C
instructions instead of the product's bit arithmetic).

Pins, independent of both restatements:
* the reference's own table tests (src/quant_weights.rs:3150-3527): the
  IDENTITY / DCT2X2 library values, the AFV fixed positions and their 4x8 /
  4x4 relations, the stored DCT8 / DCT16X8 tables, the DCT16X16 DC weights,
  forward x inverse = 1, the new_fast B base weight of 512;
* an FP64 evaluation of the format's radial band formula (GetQuantWeights:
  bands b_0 = p_0, b_i = b_{i-1} * mult(p_i), geometric interpolation at
  radius sqrt(dx^2 + dy^2) scaled to num_bands - 1) fed with the reference's
  band constants, each f16-rounded with numpy's float16 (round to nearest
  even) where the reference signals them, checked to 2e-5: a wrong tier, B
  row, SS2 scale, tilt, f16 rule, table shape or transposition fails it;
* the tier boundaries of new_fast (0.3, 1.25, qm_ss2_distance 2.25 / 1.75,
  3.5), from the constants at src/quant_weights.rs:104-121 and
  src/arbiter.rs:50-62.
"""
import numpy as np
import pytest

import jixel_amd as J

DISTANCES = [0.25, 1.0, 1.6, 2.35, 4.0]

# reference band constants (src/quant_weights.rs:123-138, 927-1230)
DCT8 = [[3150.0, 0.0, -0.4, -0.4, -0.4, -2.0], [560.0, 0.0, -0.3, -0.3, -0.3, -0.3],
        [512.0, -2.0, -1.0, 0.0, -1.0, -2.0]]
FLAT_HQ = [512.0, -0.5, -0.25, 0.0, -0.25, -0.5]
FLAT_MID = [512.0, -0.13, -0.235, -0.276, -0.609, -0.339]
DCT16 = [[8996.872571181412, -1.3000777393353804, -0.49424529824571225, -0.43909377445710344,
          -0.6350101832695744, -0.9017726405082761, -1.6162099239887414],
         [3191.4836629684423, -0.67424582104194355, -0.80745813428471, -0.4492583748484344,
          -0.35865440981033403, -0.31322389111877305, -0.37615025315725483],
         [1157.504081454872, -2.0531423165804414, -1.4, -0.506871300333784,
          -0.42708730624733904, -1.4856834539296244, -4.92091428844016]]
DCT32 = [[15718.40830982519, -1.025, -0.98, -0.9012, -0.4, -0.48819395, -0.421064, -0.27],
         [7305.763681069598, -0.80419582, -0.76330365, -0.5566038, -0.49785305, -0.43699593,
          -0.40180867, -0.27321683],
         [3803.53173721215, -3.0607336, -2.041327, -2.023565, -0.54953895, -0.4, -0.4, -0.3]]
DCT16X32 = [[13844.97076442300573, -0.97113799999999995, -0.658, -0.42026, -0.22712, -0.2206,
             -0.226, -0.6],
            [4798.964084220744293, -0.61125308982767057, -0.83770786552491361,
             -0.79014862079498627, -0.2692727459704829, -0.38272769465388551,
             -0.22924222653091453, -0.20719098826199578],
            [1807.236946760964614, -1.2, -1.2, -0.7, -0.7, -0.7, -0.4, -0.5]]
DCT16X8 = [[7240.7734393502, -0.7, -0.7, -0.2, -0.2, -0.2, -0.5],
           [1448.15468787004, -0.5, -0.5, -0.5, -0.2, -0.2, -0.2],
           [506.854002029, -1.4, -0.2, -0.5, -0.5, -1.5, -3.6]]
DCT4X8 = [[2198.050556016380522, -0.96269623020744692, -0.76194253026666783, -0.6551140670773547],
          [764.3655248643528689, -0.92630200888366945, -0.9675229603596517, -0.27845290869168118],
          [527.107573587542228, -1.4594385811273854, -1.450082094097871593, -1.5843722511996204]]
DCT4X4 = [[2200.0, 0.0, 0.0, 0.0], [392.0, 0.0, 0.0, 0.0], [112.0, -0.25, -0.25, -0.5]]
AFV = [[3072.0, 3072.0, 256.0, 256.0, 256.0, 414.0, 0.0, 0.0, 0.0],
       [1024.0, 1024.0, 50.0, 50.0, 50.0, 58.0, 0.0, 0.0, 0.0],
       [384.0, 384.0, 12.0, 12.0, 12.0, 22.0, -0.25, -0.25, -0.25]]
IDENTITY = [[280.0, 3160.0, 3160.0], [60.0, 864.0, 864.0], [18.0, 200.0, 200.0]]
DCT2 = [[3840.0, 2560.0, 1280.0, 640.0, 480.0, 300.0], [960.0, 640.0, 320.0, 180.0, 140.0, 120.0],
        [640.0, 320.0, 128.0, 64.0, 32.0, 16.0]]

K = dict(DCT8=0, DCT16X8=1, DCT16X16=2, DCT32X32=3, DCT32X16=4, DCT8X32=5, DCT64X64=6,
         DCT64X32=7, IDENTITY=8, DCT2X2=9, DCT4X4=10, DCT4X8=11, AFV=12)


def f16(v):
    return float(np.float16(np.float32(v)))


def wire_abs(v):
    return f16(np.float32(v) / np.float32(64.0)) * 64.0


def mult(v):
    return 1.0 + v if v > 0 else 1.0 / (1.0 - v)


def band_matrix64(params, rows, cols):
    """FP64 GetQuantWeights: step sizes 1 / weight, rows x cols."""
    nb = len(params)
    bands = [params[0]]
    for v in params[1:]:
        bands.append(bands[-1] * mult(v))
    scale = (nb - 1) / (np.sqrt(2.0) + 1e-6)
    out = np.zeros((rows, cols))
    for y in range(rows):
        for x in range(cols):
            d = np.hypot(x * scale / (cols - 1), y * scale / (rows - 1))
            i = min(int(np.floor(d)), nb - 2)
            out[y, x] = 1.0 / (bands[i] * (bands[i + 1] / bands[i]) ** (d - i))
    return out


def mat(blob, kind, c):
    n = J.qt_size(K[kind]) if isinstance(kind, str) else J.qt_size(kind)
    off = J.qt_offset(K[kind])
    return blob[off + c * n: off + (c + 1) * n]


def inv(blob, kind, c):
    n = J.qt_size(K[kind])
    off = J.qt_offset(K[kind])
    return blob[off + (3 + c) * n: off + (4 + c) * n]


def tier(d, arbiter=0):
    ss2 = 1.75 if arbiter == 1 else 2.25
    if d < 0.3:
        return 0
    if d >= ss2:
        return 4 if d >= 3.5 else 3
    return 2 if d >= 1.25 else 1


def expected_dct8_bands(d):
    t = tier(d)
    if t == 0:
        return None
    rows = [list(r) for r in DCT8]
    if t == 4:  # coarse_dct8_override: luma steps tilted by 1.25^(1/5)
        step = 1.25 ** 0.2
        for i in range(1, 6):
            r = mult(DCT8[1][i]) * step
            rows[1][i] = r - 1.0 if r > 1.0 else 1.0 - 1.0 / r
    rows[2] = list(FLAT_HQ if t == 1 else FLAT_MID)
    return [[wire_abs(v) if i == 0 else f16(v) for i, v in enumerate(r)] for r in rows]


def scaled_bands(rows, s):
    return [[wire_abs(np.float32(r[0]) * np.float32(s))] + [f16(v) for v in r[1:]] for r in rows]


@pytest.fixture(scope="module", params=["product", "oracle"])
def tables(request, oracle):
    if request.param == "product":
        return lambda d, fused=True, arbiter=0: J.default_quant_tables(d, fused, arbiter)
    return lambda d, fused=True, arbiter=0: oracle.fast_tables(d, fused, arbiter)


def test_restatements_agree_bit_for_bit(oracle):
    ds = [0.05, 0.25, 0.299, 0.3, 0.74, 1.0, 1.2499, 1.25, 1.6, 1.75, 2.0, 2.2499, 2.25, 2.35,
          3.4999, 3.5, 4.0, 4.6, 9.0, 25.0]
    for d in ds:
        for fused in (True, False):
            for arb in (0, 1):
                a = J.default_quant_tables(d, fused, arb).view(np.uint32)
                b = oracle.fast_tables(d, fused, arb).view(np.uint32)
                assert np.array_equal(a, b), (d, fused, arb, np.flatnonzero(a != b)[:8])


@pytest.mark.parametrize("d", DISTANCES)
@pytest.mark.parametrize("fused", [True, False])
def test_band_tables_follow_the_reference_constants(tables, golden, d, fused):
    blob = tables(d, fused)
    b8 = expected_dct8_bands(d)
    for c in range(3):
        if b8 is None:  # the stored library table (spec tier, new(d) for d < 0.3)
            assert np.array_equal(mat(blob, "DCT8", c), np.float32(golden["DEQUANT_MATRIX_8X8"][c]))
        else:
            np.testing.assert_allclose(mat(blob, "DCT8", c),
                                       band_matrix64(b8[c], 8, 8).ravel(), rtol=2e-5)
    ss2 = tier(d) >= 3
    for kind, rows, shape, s in [("DCT16X16", DCT16, (16, 16), 0.78), ("DCT32X32", DCT32, (32, 32), 0.89),
                                 ("DCT32X16", DCT16X32, (16, 32), 1.20)]:
        bands = scaled_bands(rows, s) if ss2 else rows
        for c in range(3):
            np.testing.assert_allclose(mat(blob, kind, c), band_matrix64(bands[c], *shape).ravel(),
                                       rtol=2e-5, err_msg=kind)
    for c in range(3):
        # the stored DCT16X8 table, 8 rows x 16 columns
        assert np.array_equal(mat(blob, "DCT16X8", c), np.float32(golden["DEQUANT_MATRIX_16X8"][c]))
        np.testing.assert_allclose(mat(blob, "DCT16X8", c), band_matrix64(DCT16X8[c], 8, 16).ravel(),
                                   rtol=4e-6)
        # DCT4X4: 4x4 weights, each on a 2x2 cell
        w44 = band_matrix64(DCT4X4[c], 4, 4)
        np.testing.assert_allclose(mat(blob, "DCT4X4", c),
                                   np.kron(w44, np.ones((2, 2))).ravel(), rtol=2e-5)
        # DCT4X8: the refit (QM_4X8_SCALE, src/quant_weights.rs:1807-1836)
        p48 = [wire_abs(np.float32(DCT4X8[c][0]) * np.float32([1.2, 1.0, 0.45][c]))] + \
              [f16(v) for v in DCT4X8[c][1:]]
        np.testing.assert_allclose(mat(blob, "DCT4X8", c),
                                   np.repeat(band_matrix64(p48, 4, 8), 2, axis=0).ravel(), rtol=2e-5)


def test_dct8x32_is_derived_from_the_pair_table(tables):
    """derive_8x32 (src/quant_weights.rs:1729-1789): the 16x8 band curve at
    the 8-band nodes, x sqrt(2), signalled in f16; here in FP64 from the
    reference constants with numpy's f16."""
    blob = tables(2.35)
    for c in range(3):
        curve = [DCT16X8[c][0]]
        for v in DCT16X8[c][1:]:
            curve.append(curve[-1] * mult(v))

        def sample(pos):
            if pos >= 6:
                return curve[6]
            i = int(np.floor(pos))
            return curve[i] * (curve[i + 1] / curve[i]) ** (pos - i)
        nodes = [sample(i * 6.0 / 7.0) * np.sqrt(2.0) for i in range(8)]
        params = [wire_abs(nodes[0])]
        for i in range(1, 8):
            r = nodes[i] / nodes[i - 1]
            params.append(f16(r - 1.0 if r >= 1.0 else 1.0 - 1.0 / r))
        np.testing.assert_allclose(mat(blob, "DCT8X32", c), band_matrix64(params, 8, 32).ravel(),
                                   rtol=2e-5)
        # at sqrt(2) the long-axis steps match the pair table's (equal error per pixel)
        np.testing.assert_allclose(mat(blob, "DCT8X32", c)[0] * np.sqrt(2.0),
                                   mat(blob, "DCT16X8", c)[0], rtol=2e-3)


@pytest.mark.parametrize("d", DISTANCES)
def test_reference_known_answers(tables, golden, d):
    blob = tables(d)
    # identity_and_dct2_tables_match_the_jpeg_xl_library (src/quant_weights.rs:3293)
    idn = inv(blob, "IDENTITY", 0)
    assert (idn[0], idn[1], idn[8], idn[9], idn[2]) == (0.0, 3160.0, 3160.0, 3160.0, 280.0)
    assert mat(blob, "IDENTITY", 0)[2] == np.float32(1.0) / np.float32(280.0)
    d2 = inv(blob, "DCT2X2", 2)
    assert (d2[0], d2[1], d2[9], d2[2], d2[18], d2[4], d2[36]) == (0, 640, 320, 128, 64, 32, 16)
    assert mat(blob, "DCT2X2", 2)[36] == np.float32(1.0 / 16.0)
    for c in range(3):
        # afv_matrix_fixed_positions_match_library_weights (src/quant_weights.rs:3417)
        mc = mat(blob, "AFV", c)
        for pos, i in [(8, 0), (1, 1), (16, 2), (2, 3), (18, 4)]:
            assert mc[pos] == np.float32(1.0) / np.float32(AFV[c][i])
        lib48 = np.repeat(band_matrix64(DCT4X8[c], 4, 8), 2, axis=0).ravel()
        np.testing.assert_allclose(mc[9:16], lib48[1:8], rtol=2e-5)  # library 4x8, not the refit
        m44 = mat(blob, "DCT4X4", c)
        for x in range(1, 4):
            assert mc[2 * x + 1] == m44[2 * x]
        assert inv(blob, "AFV", c)[0] == 0.0 and np.isfinite(mc).all() and (mc > 0).all()
        # dct16x16_matrix_dc_matches_libjxl_polynomial (src/quant_weights.rs:3477), spec tiers
        if tier(d) < 3:
            assert abs(mat(blob, "DCT16X16", c)[0] - 1.0 / DCT16[c][0]) < 1e-7
        m16 = mat(blob, "DCT16X16", c)
        assert (m16 > 0).all() and m16[255] > m16[0]
    # ordinary_dct8_fit_matches_wire_weights_and_existing_gates (:3331): new_fast keeps
    # the B base weight 512 (no ordinary fit)
    if tier(d) > 0:
        assert mat(blob, "DCT8", 2)[0] == np.float32(1.0 / 512.0)
    # forward x inverse (src/quant_weights.rs:3150-3174)
    for kind in K:
        for c in range(3):
            m, v = mat(blob, kind, c), inv(blob, kind, c)
            assert v[0] == 0.0
            assert np.abs(m[1:].astype(np.float64) * v[1:] - 1.0).max() <= 4 * np.finfo(np.float32).eps


def test_tier_boundaries(tables):
    spec8 = tables(0.1)
    for d, want in [(0.299, 0), (0.3, 1), (1.2499, 1), (1.25, 2), (2.2499, 2), (2.25, 3),
                    (3.4999, 3), (3.5, 4), (9.0, 4)]:
        blob = tables(d)
        same = [np.array_equal(blob, tables(x)) for x in (0.1, 0.5, 1.5, 3.0, 4.0)]
        assert same.index(True) == want, (d, same)
    # CVVDP moves the SS2 gate to 1.75 (src/arbiter.rs:60)
    assert np.array_equal(tables(1.8, True, 1), tables(3.0, True, 0))
    assert np.array_equal(tables(1.7, True, 1), tables(1.5, True, 0))
    assert not np.array_equal(spec8, tables(0.5))
