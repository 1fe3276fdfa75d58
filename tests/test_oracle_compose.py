"""Pins of the composed quantize stage: CfL subtraction, the Y dequant round
trip and the DC outputs (src/group.rs:795-818,1295-1306,1840-1870,
2000-2030; src/color_correlation.rs:1117-1125,1597-1610).

Single-level cases: Y, X and B blocks are each the FP64 inverse DCT8 of one
coefficient at the same position k (and a pixel offset, which is the DCT8 DC).
Then, in closed form from the reference's definitions,

* Y's level at k is L when a = c_Y * inv_y[k] * qac = L + 1/4;
* the dequantized Y is Ydq = level(L) * qm_y[k] / qac, with level(+-1) =
  +-one and level(L) = L - numerator / L otherwise (DequantBias::level);
* the stored X and B coefficients are c_X - x_factor * Ydq and
  c_B - b_factor * Ydq, with x_factor = base_x + ytox / 84 and b_factor =
  base_b + ytob / 84 of the tile's map entries, and are 0 elsewhere
  (Y's lowest frequency is zeroed before the subtraction, so X's and B's
  DC coefficients keep their own value);
* X and B levels at k are the chosen targets when c_X and c_B are placed so
  that the CfL-subtracted coefficient times inv_qm * qac * qm_mul is
  target + 1/4;
* the DC levels are round(dc * 4096/dc_step[0] * scale_dc) for X (base_x
  = 0 selects the quantizer without CfL), round(dc * 512/dc_step[1] *
  scale_dc) for Y, and round(dc_B * 256/dc_step[2] * scale_dc - ydc_q *
  256/dc_step[2] * dc_step[1]/512 * (base_b + ytob_dc / 84)) for B.

Proportional chroma: block-constant planes with B = b_slope * Y, where the
B DC prediction from the stored Y DC leaves exactly the Y DC's rounding
residue, and a DCT16X16 region whose four DC values are the region's value.

Using the source Y instead of the dequantized one, a sign error in the
subtraction, the B slope without ytob_dc, the wrong DC scale or the zeroed
LLF left in fails these."""
import math

import numpy as np
import pytest

from test_oracle_single_coeff import block_for


@pytest.fixture(scope="module")
def J():
    import jixel_amd
    return jixel_amd


def params(J, distance=1.0):
    return J.frame_params(J.EncodeConfig().with_distance(distance))


def dct8_tables(J, tables):
    off, n = J.qt_offset(0), 64
    qm = [np.asarray(tables[off + c * n: off + (c + 1) * n], np.float64) for c in range(3)]
    inv = [np.asarray(tables[off + (3 + c) * n: off + (4 + c) * n], np.float64) for c in range(3)]
    return qm, inv


def level_value(p, L):
    """DequantBias::level for channel Y (src/group.rs:770-790)."""
    if L == 0:
        return 0.0
    if abs(L) == 1:
        return math.copysign(float(p.dequant_one[1]), L)
    return L - float(p.dequant_numerator) / L


def qm_muls(p):
    return 1.25 ** (p.x_qm_scale - 2.0), 1.25 ** (min(max(p.b_qm_scale, 2), 7) - 2.0)


def test_tables_are_reciprocal(J, oracle):
    p = params(J)
    qm, inv = dct8_tables(J, oracle.tables_for(p))
    for c in range(3):
        # the DC position's inverse weight is 0: the LLF is not AC-quantized
        assert inv[c][0] == 0.0
        assert np.allclose(qm[c][1:] * inv[c][1:], 1.0, rtol=1e-6)


CASES = [  # (k, Y level, ytox, ytob, X target, B target)
    (9, 3, 12, -20, 2, -3),
    (1, 1, -30, 5, 4, 2),
    (27, -2, 40, 0, -5, 6),
    (36, 5, 0, -84, 3, -2),
    (8, -1, 7, 33, -2, 2),
    (63, 2, -12, -40, 2, 3),
]


def cfl_inputs(J, oracle, case, fused):
    """Inputs of one single-level case and the closed-form values."""
    k, L, ytox_v, ytob_v, tx_level, tb_level = case
    oracle.set_fused(fused)
    p = params(J)
    p.fused = 1 if fused else 0
    p.x_qm_scale, p.b_qm_scale = (3, 4) if k % 2 else (2, 2)
    tables = oracle.tables_for(p)
    qm, inv = dct8_tables(J, tables)
    x_mul, b_mul = qm_muls(p)
    qfv = 9
    qac = float(np.float32(p.scale) * np.float32(qfv))
    u, v = divmod(k, 8)
    cf = float(p.cfl_color_factor)
    x_factor = float(p.cfl_base_x) + ytox_v / cf
    b_factor = float(p.cfl_base_b) + ytob_v / cf
    c_y = (L + math.copysign(0.25, L)) / (inv[1][k] * qac)
    ydq = level_value(p, L) * qm[1][k] / qac
    x_post = (tx_level + math.copysign(0.25, tx_level)) / (inv[0][k] * qac * x_mul)
    b_post = (tb_level + math.copysign(0.25, tb_level)) / (inv[2][k] * qac * b_mul)
    c_x = x_post + x_factor * ydq
    c_b = b_post + b_factor * ydq
    xyb = np.zeros((3, 64, 64), np.float32)
    sl = (slice(16, 24), slice(8, 16))  # block (1, 2)
    xyb[1][sl] = block_for(u, v, c_y)
    xyb[0][sl] = block_for(u, v, c_x)
    xyb[2][sl] = block_for(u, v, c_b)
    return dict(p=p, tables=tables, xyb=xyb, cells=np.full((8, 8), 1, np.uint8),
                qf=np.full((8, 8), qfv, np.uint8), ytox=np.full((1, 1), ytox_v, np.int8),
                ytob=np.full((1, 1), ytob_v, np.int8), c_y=c_y, x_post=x_post, b_post=b_post)


def run(oracle, a):
    return oracle.forward_quantize(a["xyb"], a["cells"], a["qf"], a["ytox"], a["ytob"],
                                   a["tables"], a["p"])


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("case", CASES)
def test_cfl_subtracts_the_dequantized_y(J, oracle, case, fused):
    k, L, _, _, tx_level, tb_level = case
    a = cfl_inputs(J, oracle, case, fused)
    c_y, x_post, b_post = a["c_y"], a["x_post"], a["b_post"]
    levels, coeffs, dc = run(oracle, a)
    blk = 2 * 8 + 1
    rng = slice(blk * 64, (blk + 1) * 64)
    for c, target in ((0, tx_level), (1, L), (2, tb_level)):
        want = np.zeros(64, np.int32)
        want[k] = target
        assert np.array_equal(levels[c][rng], want), (c, levels[c][rng][k], target)
        others = np.delete(levels[c].reshape(-1, 64), blk, axis=0)
        assert not others.any(), c
    # the stored coefficients: Y is the source, X and B are CfL-subtracted
    for c, want_k in ((1, c_y), (0, x_post), (2, b_post)):
        got = coeffs[c][rng].astype(np.float64)
        scale = max(abs(want_k), abs(c_y))
        assert abs(got[k] - want_k) <= 2e-6 * scale, (c, got[k], want_k)
        rest = np.delete(got, k)
        assert np.abs(rest).max() <= 2e-6 * scale, (c, np.abs(rest).max())
    assert not any(d.any() for d in dc)


def block16_for(coeffs):
    """16x16 pixels whose DCT16X16, in the reference's square storage
    (index u * 16 + v = vertical frequency v, horizontal frequency u), is
    `coeffs` ({index: value})."""
    n = 16
    k = np.arange(n)[:, None]
    i = np.arange(n)[None, :]
    m = np.cos(np.pi * k * (2 * i + 1) / (2 * n))
    m[1:] *= math.sqrt(2.0)
    ref = np.zeros((n, n))
    for idx, c in coeffs.items():
        u, v = divmod(idx, n)
        ref[v, u] = c
    inv = np.linalg.inv(m)
    return (n * n) * inv @ ref @ inv.T


def llf_inputs(J, oracle):
    oracle.set_fused(True)
    p = params(J)
    tables = oracle.tables_for(p)
    S = J.STRATEGY
    kind = oracle.lib().jxo_qt_kind(S["DCT16X16"])
    off, n = J.qt_offset(kind), 256
    qm_y = np.asarray(tables[off + n: off + 2 * n], np.float64)
    inv_y = np.asarray(tables[off + 4 * n: off + 5 * n], np.float64)
    llf, ac = 1, 34
    assert inv_y[llf] > 0 and inv_y[ac] > 0
    qfv = 9
    qac = float(np.float32(p.scale) * np.float32(qfv))
    c_llf = 3.25 / (inv_y[llf] * qac)
    c_ac = 2.25 / (inv_y[ac] * qac)
    ytox_v = 20
    xyb = np.zeros((3, 64, 64), np.float32)
    xyb[1, 32:48, 16:32] = block16_for({llf: c_llf, ac: c_ac})
    cells = np.full((8, 8), (S["DCT"] << 1) | 1, np.uint8)
    cells[4:6, 2:4] = S["DCT16X16"] << 1
    cells[4, 2] = (S["DCT16X16"] << 1) | 1
    return dict(p=p, tables=tables, xyb=xyb, cells=cells, qf=np.full((8, 8), qfv, np.uint8),
                ytox=np.full((1, 1), ytox_v, np.int8), ytob=np.full((1, 1), -84, np.int8),
                llf=llf, ac=ac, c_llf=c_llf, c_ac=c_ac, qac=qac, qm_y=qm_y,
                x_factor=float(p.cfl_base_x) + ytox_v / float(p.cfl_color_factor))


def test_cfl_uses_y_with_its_lowest_frequencies_zeroed(J, oracle):
    """A DCT16X16 whose Y has a quantized coefficient at a lowest-frequency
    position (storage 1, one of the 2x2 LLF) and one at an AC position: the
    CfL-subtracted X keeps 0 at the LLF position (Y's LLF is zeroed before the
    subtraction) and is -x_factor * Ydq at the AC position."""
    a = llf_inputs(J, oracle)
    p, llf, ac, c_llf, c_ac = a["p"], a["llf"], a["ac"], a["c_llf"], a["c_ac"]
    levels, coeffs, _ = run(oracle, a)

    def gather(a):
        return np.concatenate([a[(by * 8 + 2) * 64:(by * 8 + 2) * 64 + 128] for by in (4, 5)])

    ly = gather(levels[1])
    assert ly[llf] == 3 and ly[ac] == 2 and np.count_nonzero(ly) == 2
    x = gather(coeffs[0]).astype(np.float64)
    ydq = level_value(p, 2) * a["qm_y"][ac] / a["qac"]
    want = -a["x_factor"] * ydq
    assert abs(x[llf]) < 1e-7 * abs(c_llf), x[llf]
    assert abs(x[ac] - want) <= 2e-6 * abs(ydq), (x[ac], want)
    assert np.abs(np.delete(x, [llf, ac])).max() < 2e-6 * abs(c_ac)


def dc_scales(p):
    sdc = float(p.scale_dc)
    ds = [float(p.dc_step[c]) for c in range(3)]
    inv_factor = [4096.0 / ds[0] * sdc, 512.0 / ds[1] * sdc, 256.0 / ds[2] * sdc]
    b_slope = float(p.cfl_base_b) + p.ytob_dc / float(p.cfl_color_factor)
    cfl_b = 256.0 / ds[2] * (ds[1] / 512.0) * b_slope
    cfl_x = 4096.0 / ds[0] * (ds[1] / 512.0) * float(p.cfl_base_x)
    return inv_factor, cfl_b, cfl_x, b_slope


def round_half_away(x):
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def dc_inputs(J, oracle, ytob_dc):
    """B = b_slope * Y + e per block, with e set so that the B DC before
    rounding is k +- 0.4 given the stored (rounded) Y DC."""
    oracle.set_fused(True)
    p = params(J)
    p.ytob_dc = ytob_dc
    tables = oracle.tables_for(p)
    inv_factor, cfl_b, cfl_x, b_slope = dc_scales(p)
    assert cfl_x == 0.0 and cfl_b > 0.3
    S = J.STRATEGY
    xb = yb = 8
    rng = np.random.default_rng(100 + ytob_dc)
    yv = rng.uniform(0.05, 0.8, (yb, xb)).astype(np.float32)
    yv[4:6, 2:4] = yv[4, 2]  # one DCT16X16 region, constant
    xv = (np.float32(0.021) * yv - np.float32(0.004)).astype(np.float32)
    bv = np.zeros_like(yv)
    want, flips = {}, 0
    for by in range(yb):
        for bx in range(xb):
            y = float(yv[by, bx])
            ydc = round_half_away(y * inv_factor[1])
            r = y * inv_factor[1] - ydc
            k = int(rng.integers(-3, 4))
            target = k + (0.4 if r < 0 else -0.4)
            b = np.float32((target + ydc * cfl_b) / inv_factor[2])
            bv[by, bx] = b
            pre_b = float(b) * inv_factor[2] - ydc * cfl_b
            pre_x = float(xv[by, bx]) * inv_factor[0]
            for v in (pre_x, pre_b):
                assert abs(v - math.floor(v) - 0.5) > 1e-3, "away from ties"
            want[by, bx] = (round_half_away(pre_x), ydc, round_half_away(pre_b))
            unrounded = float(b) * inv_factor[2] - y * inv_factor[1] * cfl_b
            flips += round_half_away(unrounded) != want[by, bx][2]
    bv[4:6, 2:4] = bv[4, 2]
    for by in (4, 5):
        for bx in (2, 3):
            want[by, bx] = want[4, 2]
    xyb = np.stack([np.kron(a, np.ones((8, 8), np.float32)) for a in (xv, yv, bv)])
    cells = np.full((yb, xb), (S["DCT"] << 1) | 1, np.uint8)
    cells[4:6, 2:4] = S["DCT16X16"] << 1
    cells[4, 2] = (S["DCT16X16"] << 1) | 1
    return dict(p=p, tables=tables, xyb=xyb, cells=cells, qf=np.full((yb, xb), 7, np.uint8),
                ytox=np.zeros((1, 1), np.int8), ytob=np.zeros((1, 1), np.int8), want=want,
                flips=flips)


@pytest.mark.parametrize("ytob_dc", [0, 10, -25])
def test_dc_levels_of_proportional_chroma(J, oracle, ytob_dc):
    """The prediction from the unrounded Y DC would move the B DC by
    cfl_b * r, r the Y DC's rounding residue, and so land on another level
    wherever cfl_b * |r| > 0.1."""
    a = dc_inputs(J, oracle, ytob_dc)
    levels, _, dc = run(oracle, a)
    assert not any(l.any() for l in levels), "constant blocks have no AC"
    dc = [d.reshape(8, 8) for d in dc]
    for (by, bx), w in a["want"].items():
        got = tuple(int(d[by, bx]) for d in dc)
        assert got == w, (by, bx, got, w)
    assert a["flips"] >= 10, a["flips"]


def all_inputs(J, oracle):
    """Every composed case, for running the same inputs elsewhere."""
    for case in CASES:
        for fused in (True, False):
            yield "cfl%s_%s" % (case[:2], fused), cfl_inputs(J, oracle, case, fused)
    yield "llf", llf_inputs(J, oracle)
    for ytob_dc in (0, 10, -25):
        yield "dc_%d" % ytob_dc, dc_inputs(J, oracle, ytob_dc)
