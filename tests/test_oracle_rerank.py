"""Pins of the oracle's SSIM rerank at the Fast tier (find_rerank_downgrades
and best_reconstruction_plan, src/ac_strategy/selection.rs:2636-3232).

* best_reconstruction_plan against an exhaustive search over every tiling
  on the reference's own seeded cases (:3884-3962: its LCG, footprints 1x2
  .. 4x4, tiles 1x1 .. 4x2), costs equal and the plan covering the
  footprint exactly;
* its known answer that keeps one child and splits the other (:3964-3996),
  and first-candidate tie keeping;
* the pass on a flat image, where every candidate costs lambda * RERANK_META_R
  (zero distortion and rate), keeps every merge (the merge pays the charge
  once, its tiling once per piece);
* on textured content the pass only splits merges, into DCT8 and aligned
  16x16 children that tile the merge's footprint;
* the gates: no change above d = 1.709, at Fastest, or with the CVVDP
  arbiter's zero margin."""
import numpy as np
import pytest

NO = 255
# covered blocks per raw strategy code (src/dc_group_data.rs:152-170)
COV_X = [1, 1, 2, 2, 4, 1, 1, 1, 2, 4, 1, 1, 1, 1, 8, 4, 8, 1, 1, 16, 8, 16, 1, 4]
COV_Y = [1, 2, 1, 2, 4, 1, 1, 1, 4, 2, 1, 1, 1, 1, 8, 8, 4, 1, 1, 16, 16, 8, 4, 1]


def lcg_tiles():
    """The reference test's cases, in its order."""
    state = 31
    cases = []
    for width, height in [(1, 2), (2, 1), (2, 2), (2, 4), (4, 2), (4, 4)]:
        for _ in range(12):
            tiles = []
            for w, h in [(1, 1), (1, 2), (2, 1), (2, 2), (2, 4), (4, 2)]:
                if w > width or h > height:
                    continue
                for y in range(0, height, h):
                    for x in range(0, width, w):
                        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
                        tiles.append((x, y, w, h, float(1 + (state >> 24))))
            cases.append((width, height, tiles))
    return cases


def mask_of(t):
    x, y, w, h, _ = t
    m = 0
    for yy in range(y, y + h):
        for xx in range(x, x + w):
            m |= 1 << (yy * 4 + xx)
    return m


def exhaustive(tiles, full):
    masks = [mask_of(t) for t in tiles]
    memo = {}

    def best(covered):
        if covered == full:
            return 0.0
        if covered in memo:
            return memo[covered]
        rest = full & ~covered
        first = rest & -rest
        b = float("inf")
        for t, m in zip(tiles, masks):
            if m & first and not (m & covered):
                b = min(b, t[4] + best(covered | m))
        memo[covered] = b
        return b

    return best(0)


def test_plan_matches_exhaustive_on_the_reference_seeded_cases(oracle):
    for width, height, tiles in lcg_tiles():
        full = 0
        for t in tiles:
            full |= mask_of(t)
        want = exhaustive(tiles, full)
        cost, grid = oracle.rerank_plan(tiles, width, height)
        assert cost == want, (width, height)
        covered = 0
        for i in grid:
            if i == NO:
                continue
            m = mask_of(tiles[i])
            assert covered & m == 0
            covered |= m
        assert covered == full


def test_plan_keeps_one_child_and_splits_the_other(oracle):
    tiles = [(x, y, 1, 1, 5.0) for y in range(2) for x in range(2)]
    tiles += [(0, 0, 1, 2, 3.0), (1, 0, 1, 2, 15.0)]
    cost, grid = oracle.rerank_plan(tiles, 2, 2)
    assert cost == 13.0
    assert grid[0] == 4 and grid[1] == 1 and grid[4] == NO and grid[5] == 3


def test_plan_ties_keep_the_first_candidate(oracle):
    # two equal 2x2 tiles: the first listed survives
    tiles = [(0, 0, 2, 2, 4.0), (0, 0, 2, 2, 4.0)]
    cost, grid = oracle.rerank_plan(tiles, 2, 2)
    assert cost == 4.0 and grid[0] == 0
    # a horizontal and a vertical split of equal cost: the horizontal halves
    # (tried first) are kept
    tiles = [(0, 0, 1, 2, 2.0), (1, 0, 1, 2, 2.0), (0, 0, 2, 1, 2.0), (0, 1, 2, 1, 2.0)]
    cost, grid = oracle.rerank_plan(tiles, 2, 2)
    assert cost == 4.0 and grid[0] == 0 and grid[1] == 1


def layout(J, xb, yb, pattern):
    """A layout of 32x32 merges (pattern 32), 16x16 merges (16), 16x8 / 8x16
    pairs (pairs) or DCT8."""
    S = J.STRATEGY
    cells = np.full((yb, xb), (S["DCT"] << 1) | 1, np.uint8)

    def put(bx, by, s, cx, cy):
        for iy in range(cy):
            for ix in range(cx):
                cells[by + iy, bx + ix] = (s << 1) | (1 if ix == 0 and iy == 0 else 0)

    for by in range(0, yb - 3, 4):
        for bx in range(0, xb - 3, 4):
            if pattern == "32":
                put(bx, by, S["DCT32X32"], 4, 4)
            elif pattern == "16":
                for o in range(4):
                    put(bx + 2 * (o & 1), by + 2 * (o >> 1), S["DCT16X16"], 2, 2)
            else:
                put(bx, by, S["DCT8X16"], 2, 1)
                put(bx + 2, by, S["DCT16X8"], 1, 2)
    return cells


def inputs(J, oracle, content, w, h, d, speed="FAST", arbiter=0):
    from jixel_amd import synth
    if content == "flat":
        rgb = np.full((h, w, 3), 117, np.uint8)
    else:
        rgb = synth.photo_like(w, h, seed=5)
    xyb = oracle.xyb_image(rgb, J.XYB_SPEC_FWD)
    cfg = J.EncodeConfig(speed=J.Speed[speed]).with_distance(d)
    p = J.frame_params(cfg)
    p.arbiter = arbiter
    qf = oracle.fill_quant_field(xyb, p.distance, 1.0 / p.scale)
    tiles = ((h + 63) // 64, (w + 63) // 64)
    z = np.zeros(tiles, np.int8)
    return xyb, p, qf, z, oracle.tables_for(p)


@pytest.mark.parametrize("pattern", ["32", "16", "pairs"])
def test_flat_image_keeps_every_merge(J, oracle, pattern):
    oracle.set_fused(True)
    xyb, p, qf, z, tables = inputs(J, oracle, "flat", 128, 96, 1.2)
    cells = layout(J, 16, 12, pattern)
    out = oracle.rerank_pass(xyb, qf, z, z, tables, p, cells)
    assert (out == cells).all()


@pytest.fixture(scope="module")
def J():
    import jixel_amd
    return jixel_amd


def covers(J, cells):
    """Every block is covered by exactly one first block's footprint."""
    yb, xb = cells.shape
    cov = np.zeros_like(cells, np.int32)
    for by in range(yb):
        for bx in range(xb):
            if cells[by, bx] & 1:
                s = cells[by, bx] >> 1
                cx, cy = COV_X[s], COV_Y[s]
                cov[by:by + cy, bx:bx + cx] += 1
    return (cov == 1).all()


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("pattern", ["32", "16", "pairs"])
def test_textured_merges_split_into_footprint_children(J, oracle, pattern, fused):
    oracle.set_fused(fused)
    xyb, p, qf, z, tables = inputs(J, oracle, "photo", 128, 96, 1.0)
    p.fused = 1 if fused else 0
    cells = layout(J, 16, 12, pattern)
    out = oracle.rerank_pass(xyb, qf, z, z, tables, p, cells)
    S = J.STRATEGY
    assert covers(J, out)
    changed = out != cells
    assert changed.any()  # photo-like detail at d = 1 splits some merges
    allowed = {S["DCT"], S["DCT16X16"]} if pattern == "32" else {S["DCT"]}
    for by, bx in zip(*np.nonzero(changed)):
        assert (out[by, bx] >> 1) in allowed
    # a 32x32 footprint is split inside itself only
    if pattern == "32":
        for by in range(0, 12, 4):
            for bx in range(0, 16, 4):
                blk = out[by:by + 4, bx:bx + 4]
                if (blk[0, 0] >> 1) != S["DCT32X32"]:
                    assert ((blk >> 1) != S["DCT32X32"]).all()


@pytest.mark.parametrize("d,speed,arbiter", [(1.75, "FAST", 0), (2.35, "FAST", 0),
                                             (1.0, "FASTEST", 0), (1.0, "FAST", 1)])
def test_gates(J, oracle, d, speed, arbiter):
    oracle.set_fused(True)
    xyb, p, qf, z, tables = inputs(J, oracle, "photo", 128, 96, d, speed, arbiter)
    cells = layout(J, 16, 12, "32")
    out = oracle.rerank_pass(xyb, qf, z, z, tables, p, cells)
    assert (out == cells).all()
