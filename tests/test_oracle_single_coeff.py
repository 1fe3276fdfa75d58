"""Single-coefficient images: an 8x8 block that is the inverse DCT8 (FP64) of
one AC coefficient c at storage position k = u * 8 + v.  Its forward DCT8 is
that one coefficient, so the quantize stage and the candidate cost follow in
closed form:

* the Y level at k is the dead-zone quantizer's (src/group.rs:408-459) on
  a = c * inv_weight[k] * qac, with the quadrant threshold of k's row and
  column half, and every other level of every channel is 0 (X = 0, B = 0,
  and the CfL factors set to 0 through ytox = 0, ytob = -84);
* two blocks whose coefficient sits at transposed positions of one quadrant
  with the same inverse weight quantize to the same level with the same
  error, so their DCT8 costs differ only by the coefficient rate model's
  zero-run term (src/inflated_cost.rs:492-546): RD_LAMBDA * R_ZERO times the
  difference of their scan positions, which come from the reference's
  K_COEFF_ORDER_8X8 (tests/golden).

A transposed storage layout, a wrong weight index, a wrong quadrant or a
wrong scan table fails these."""
import ctypes
import math

import numpy as np
import pytest

RD_LAMBDA = np.float32(0.08086717)
R_ZERO = 0.5


def c8():
    k = np.arange(8)[:, None]
    i = np.arange(8)[None, :]
    m = np.cos(np.pi * k * (2 * i + 1) / 16)
    m[1:] *= math.sqrt(2.0)
    return m


def block_for(u, v, c):
    """8x8 pixels whose DCT8 in the reference's storage (k = u * 8 + v, the
    transposed [vertical, horizontal] frequency array) is c at (u, v)."""
    d = np.zeros((8, 8))
    d[v, u] = c  # storage[u * 8 + v] = d[v, u]
    C = c8()
    inv = np.linalg.inv(C)
    return inv @ (64.0 * d) @ inv.T


@pytest.fixture(scope="module")
def J():
    import jixel_amd
    return jixel_amd


def setup(J, oracle, distance=2.35):
    p = J.frame_params(J.EncodeConfig().with_distance(distance))
    tables = oracle.tables_for(p)
    off = J.qt_offset(0)
    inv_y = np.asarray(tables[off + 4 * 64: off + 5 * 64], np.float32)  # inv, channel 1
    return p, tables, inv_y


def thresholds(oracle, c, distance):
    thr = np.zeros(4, np.float32)
    oracle.lib().jxo_ac_thresholds(c, 1, 1, np.float32(distance), np.float32(1.0),
                                   thr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return thr


def test_the_block_has_one_coefficient(oracle):
    for (u, v) in [(0, 1), (3, 5), (7, 7), (6, 2)]:
        plane = np.zeros((16, 16), np.float32)
        plane[8:16, 8:16] = block_for(u, v, 0.75).astype(np.float32)
        coeffs = oracle.forward_transform(0, plane, 8, 8).reshape(-1)
        want = np.zeros(64)
        want[u * 8 + v] = 0.75
        assert np.abs(coeffs - want).max() < 1e-6, (u, v)


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("level", [1, 3, -2, 9])
def test_single_coefficient_levels(J, oracle, fused, level):
    oracle.set_fused(fused)
    p, tables, inv_y = setup(J, oracle)
    p.fused = 1 if fused else 0
    qfv = 12
    qac = np.float32(p.scale) * np.float32(qfv)
    thr = thresholds(oracle, 1, p.distance)
    w = h = 64
    xb = yb = 8
    cells = np.full((yb, xb), 1, np.uint8)  # DCT8 everywhere, (0 << 1) | first
    qf = np.full((yb, xb), qfv, np.uint8)
    ytox = np.zeros((1, 1), np.int8)
    ytob = np.full((1, 1), -84, np.int8)  # B factor 1 + (-84) / 84 = 0
    checked = 0
    for k in range(1, 64):
        u, v = divmod(k, 8)
        t = thr[(2 if u >= 4 else 0) + (1 if v >= 4 else 0)]
        # a = level + 0.25 (away from the rounding ties), above the threshold
        a = level + (0.25 if level > 0 else -0.25)
        if abs(a) < t * 1.01:
            continue
        c = a / (float(inv_y[k]) * float(qac))
        xyb = np.zeros((3, h, w), np.float32)
        xyb[1, 8:16, 8:16] = block_for(u, v, c).astype(np.float32)
        levels, _, dc = oracle.forward_quantize(xyb, cells, qf, ytox, ytob, tables, p)
        blk = 1 * xb + 1
        want = np.zeros(64, np.int32)
        want[k] = level
        assert np.array_equal(levels[1][blk * 64:(blk + 1) * 64], want), (k, level)
        assert not levels[0].any() and not levels[2].any(), k
        others = np.delete(levels[1].reshape(-1, 64), blk, axis=0)
        assert not others.any(), k
        checked += 1
    assert checked >= 40


def test_below_the_threshold_is_zero(J, oracle):
    oracle.set_fused(True)
    p, tables, inv_y = setup(J, oracle)
    qfv = 12
    qac = np.float32(p.scale) * np.float32(qfv)
    thr = thresholds(oracle, 1, p.distance)
    cells = np.full((8, 8), 1, np.uint8)
    qf = np.full((8, 8), qfv, np.uint8)
    for k in (9, 14, 36, 63):
        u, v = divmod(k, 8)
        t = float(thr[(2 if u >= 4 else 0) + (1 if v >= 4 else 0)])
        for frac, want in ((0.97, 0), (1.03, int(np.round(1.03 * t)) or 1)):
            a = frac * t
            c = a / (float(inv_y[k]) * float(qac))
            xyb = np.zeros((3, 64, 64), np.float32)
            xyb[1, 8:16, 8:16] = block_for(u, v, c).astype(np.float32)
            levels, _, _ = oracle.forward_quantize(xyb, cells, qf, np.zeros((1, 1), np.int8),
                                                   np.full((1, 1), -84, np.int8), tables, p)
            got = levels[1][9 * 64 + k]
            # above the threshold the level is round(a) (ties away from zero)
            assert got == (0 if frac < 1 else int(math.floor(a + 0.5))), (k, frac, got, t)


def test_cost_difference_is_the_zero_run_rate(J, oracle, golden):
    """DCT8 cost of (u, v) minus that of (v, u), same quadrant and weight:
    RD_LAMBDA * R_ZERO * (scan(u, v) - scan(v, u))."""
    oracle.set_fused(True)
    p, tables, inv_y = setup(J, oracle)
    order = np.array(golden["K_COEFF_ORDER_8X8"], np.int64)
    scan = np.empty(64, np.int64)
    scan[order] = np.arange(64)
    qac = float(np.float32(p.scale) * np.float32(10))
    thr = thresholds(oracle, 1, p.distance)
    tested = 0
    for u in range(8):
        for v in range(u + 1, 8):
            if (u >= 4) != (v >= 4):
                continue  # transposition changes the quadrant
            k1, k2 = u * 8 + v, v * 8 + u
            if inv_y[k1] != inv_y[k2] or scan[k1] == scan[k2]:
                continue
            t = float(thr[(2 if u >= 4 else 0) + (1 if v >= 4 else 0)])
            a = max(2.25, t * 1.5)
            costs = []
            for (uu, vv) in ((u, v), (v, u)):
                c = a / (float(inv_y[uu * 8 + vv]) * qac)
                xyb = np.zeros((3, 16, 16), np.float32)
                xyb[1, 8:16, 8:16] = block_for(uu, vv, c).astype(np.float32)
                costs.append(oracle.strategy_cost(xyb, J.STRATEGY["DCT"], 8, 8, qac,
                                                  [0.0, 0.0, 0.0], tables, p, meta_r=0.0))
            want = float(RD_LAMBDA) * R_ZERO * float(scan[k1] - scan[k2])
            assert abs((costs[0] - costs[1]) - want) <= 1e-5 * max(1.0, abs(costs[0])), (u, v)
            tested += 1
    assert tested >= 10
