"""Pins of the strategy search's decision pieces in the oracle, from the
reference's own tests of them (src/ac_strategy/selection.rs).

* choose_merge_partition (:115-150) against an independent footprint
  oracle over all 32 subsets of the five merges, overlaps rejected, on the
  reference's 500 seeded cases (:3782-3835), and its two known answers
  (:3838-3858, :3880-3895);
* the Fast tier's square-only decision (the oracle's choose_square32, the
  one the GPU kernel follows) equals choose_merge_partition with the
  rectangles priced +inf, on the same cases;
* aggregate_qac_2x2 against aggregate_quant (:4656-4661), the CfL factors
  of signalled tile maps (:4663-4671), and the Fast scope on a flat 32x32
  region (:4107-4150): it merges to DCT32X32 and offers no other shape."""
import math

import numpy as np
import pytest

INF = float("inf")


def lcg(state):
    while True:
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        yield float(1 + (state >> 24))


def f32(x):
    return float(np.float32(x))


def exhaustive_optimum(children, baseline, merges, accept):
    """All 32 subsets of the merges with their footprints, overlapping ones
    rejected, in f32 like the reference's test."""
    footprints = [0b1111, 0b0101, 0b1010, 0b0011, 0b1100]
    c = [np.float32(children[i // 2][i % 2]) for i in range(4)]
    b = [np.float32(baseline[i // 2][i % 2]) for i in range(4)]
    optimum = ((c[0] + c[1]) + c[2]) + c[3]
    for mask in range(1, 32):
        occupied, cost = 0, np.float32(0)
        incumbent, limit = np.float32(0), np.float32(0)
        valid = True
        for k in range(5):
            if not mask & (1 << k):
                continue
            cells = footprints[k]
            child, base = np.float32(0), np.float32(0)
            for i in range(4):
                if cells & (1 << i):
                    child = np.float32(child + c[i])
                    base = np.float32(base + b[i])
            valid &= occupied & cells == 0
            incumbent = np.float32(incumbent + child)
            limit = np.float32(limit + np.float32(base * np.float32(accept[k])))
            occupied |= cells
            cost = np.float32(cost + np.float32(merges[k]))
        valid &= bool(cost < incumbent and cost < limit)
        for i in range(4):
            if not occupied & (1 << i):
                cost = np.float32(cost + c[i])
        if valid:
            optimum = min(optimum, cost)
    return float(optimum)


def cases(n=500, seed=17):
    g = lcg(seed)
    for _ in range(n):
        children = [[next(g), next(g)], [next(g), next(g)]]
        baseline = [[v + next(g) for v in row] for row in children]
        merges = [next(g) for _ in range(5)]
        yield children, baseline, merges


def test_merge_partitions_match_exhaustive_nonoverlapping_subsets(oracle):
    accept = [0.5, 0.6, 0.7, 0.8, 0.9]
    chosen_masks = set()
    for children, baseline, merges in cases():
        chosen = oracle.choose_merge_partition(children, baseline, merges, accept)
        chosen_masks.add(chosen)
        assert oracle.partition_cost(children, merges, chosen) == \
            exhaustive_optimum(children, baseline, merges, accept)
    assert len(chosen_masks) >= 4  # the cases reach several layouts


def test_rejected_square_does_not_hide_a_single_winning_rectangle(oracle):
    children = [[25.0, 25.0], [25.0, 25.0]]
    merges = [60.0, 15.0, 90.0, 80.0, 80.0]
    assert oracle.choose_merge_partition(children, children, merges, [0.5, 0.9, 0.9, 0.9, 0.9]) == 2
    assert oracle.partition_cost(children, merges, 2) == 65.0


def test_two_rectangles_keep_their_joint_acceptance_margin(oracle):
    children = [[25.0, 25.0], [25.0, 25.0]]
    assert oracle.choose_merge_partition(children, children, [100.0, 10.0, 40.0, 100.0, 100.0],
                                         [0.7] * 5) == 6


@pytest.mark.parametrize("accept", [0.5, 0.9, 1.0, 1.2])
def test_square_scope_is_the_partition_with_rectangles_at_infinity(oracle, accept):
    agree = merged = 0
    for children, baseline, merges in cases(seed=29):
        m = merges[0] * 2.0  # passes its gates in some cases, not in others
        want = oracle.choose_merge_partition(children, baseline, [m, INF, INF, INF, INF],
                                             [accept] * 5)
        got = oracle.choose_square32(children, baseline, m, accept)
        assert got == (want == 1)
        agree += 1
        merged += got
    assert 0 < merged < agree


def test_quant_aggregation_matches_scaled_candidate_cost(oracle):
    qac = [[2.5, 5.0], [7.5, 10.0]]
    raw = oracle.aggregate_quant(40, 100, 4, 2.0)
    assert oracle.aggregate_qac_2x2(qac, 0.25, 2.0) == f32(raw * np.float32(0.25))


def test_strategy_cfl_factors_match_signaled_tile_maps(oracle):
    import jixel_amd as J
    p = J.frame_params(J.EncodeConfig())
    assert (p.cfl_base_x, p.cfl_base_b, p.cfl_color_factor) == (0.0, 1.0, 84.0)  # CflFrame::XYB
    assert list(oracle.cmap_factors(42, -42, p)) == [0.5, 0.0, 0.5]


@pytest.mark.parametrize("fused", [True, False])
def test_fast_scope_merges_a_flat_region_to_a_square(oracle, fused):
    """fill_ac_strategy at Fast on a 32x32 all-zero opsin image, quant field
    8, distance 3, global scale 1, x_qm_scale 2: DCT32X32 at (0, 0), and
    only square shapes anywhere."""
    import jixel_amd as J
    oracle.set_fused(fused)
    p = J.frame_params(J.EncodeConfig().with_distance(3.0))
    p.scale = 1.0
    p.x_qm_scale = 2
    p.fused = 1 if fused else 0
    xyb = np.zeros((3, 32, 32), np.float32)
    qf = np.full((4, 4), 8, np.uint8)
    zeros = np.zeros((1, 1), np.int8)
    cells, _ = oracle.select_strategy(xyb, qf, zeros, zeros, oracle.tables_for(p), p)
    S = J.STRATEGY
    assert cells[0, 0] >> 1 == S["DCT32X32"] and cells[0, 0] & 1
    firsts = {int(c >> 1) for c in cells.reshape(-1) if c & 1}
    assert firsts <= {S["DCT"], S["DCT16X16"], S["DCT32X32"]}
