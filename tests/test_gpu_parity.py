"""GPU parity: every device stage, called through the C ABI, against the CPU
oracle on the same seeded inputs.  Integer outputs (quant field, strategy
cells, levels, DC) and the float coefficient/cost outputs are compared bit
for bit: the kernels follow the reference's operation order exactly."""
import numpy as np
import pytest

from conftest import rand_rgb

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def J():
    import jixel_amd
    return jixel_amd


@pytest.fixture(scope="module")
def fe(J):
    f = J.FrontEnd(0)
    yield f
    f.close()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def padded(J, xyb):
    """(3, H, W) -> the device plane layout, edge-replicated."""
    _, h, w = xyb.shape
    s, ph = J.plane_geometry(w, h)
    return np.ascontiguousarray(np.pad(xyb, ((0, 0), (0, ph - h), (0, s - w)), mode="edge"))


def params_for(J, distance, fused=True, speed=None):
    cfg = J.EncodeConfig(distance=float(np.float32(distance)), fused=fused)
    if speed is not None:
        cfg = cfg.with_speed(speed)
    return J.frame_params(cfg)


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype.kind == "f":
        return np.array_equal(a.view(np.uint32 if a.itemsize == 4 else np.uint64),
                              b.view(np.uint32 if b.itemsize == 4 else np.uint64))
    return np.array_equal(a, b)


def first_mismatch(a, b):
    a, b = np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)
    if a.dtype.kind == "f":
        bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    else:
        bad = np.nonzero(a != b)[0]
    if bad.size == 0:
        return None
    i = bad[0]
    return f"{bad.size} mismatches, first at {i}: got {a[i]!r} want {b[i]!r}"


SIZES = [(1, 1), (7, 9), (65, 33), (100, 75), (256, 256)]


@pytest.mark.parametrize("order,fused", [(0, True), (0, False), (1, True)])
@pytest.mark.parametrize("w,h,ch", [(1, 1, 3), (7, 9, 4), (65, 33, 3), (300, 70, 4), (40, 30, 1),
                                    (33, 17, 2)])
def test_xyb_bit_exact(J, fe, torch, oracle, w, h, ch, order, fused):
    log("xyb", w, h, ch, order, fused)
    rgb = rand_rgb(h, w, ch, seed=w * 31 + h)
    oracle.set_fused(fused)
    want = oracle.xyb_image(rgb, J.XYB_SPEC_FWD, order)
    oracle.set_fused(True)
    got = fe.xyb(to_dev(torch, rgb), w, h, order=order, fused=fused).cpu().numpy()
    assert same_bits(got[:, :h, :w], want), first_mismatch(got[:, :h, :w], want)
    # padding is edge replication
    assert same_bits(got, padded(J, want))


@pytest.mark.parametrize("distance", [0.5, 1.0, 1.5, 2.35])
@pytest.mark.parametrize("w,h", [(27, 28), (1, 1), (7, 9), (32, 32), (33, 33), (64, 17), (65, 65),
                                 (256, 256), (2100, 40)])
def test_quant_field_bit_exact(J, fe, torch, oracle, w, h, distance):
    log("aq", w, h, distance)
    oracle.set_fused(True)
    from jixel_amd.synth import photo_like
    rgb = photo_like(w, h, seed=w + h)
    xyb = oracle.xyb_image(rgb, J.XYB_SPEC_FWD)
    p = params_for(J, distance)
    inv_scale = 1.0 / p.scale
    want = oracle.fill_quant_field(xyb, p.distance, inv_scale)
    got = fe.fill_quant_field(to_dev(torch, padded(J, xyb)), w, h, p.distance, inv_scale).cpu().numpy()
    assert same_bits(got, want), first_mismatch(got, want)


@pytest.mark.parametrize("distance", [0.5, 1.0, 1.5])
def test_quant_field_pinned_images_bit_exact(J, fe, torch, oracle, golden, distance):
    """The constant, step, blue, checkerboard and textured planes whose float
    field the oracle matches to an FP64 model (test_oracle_aq)."""
    from test_oracle_aq import cases
    p = params_for(J, distance)
    inv_scale = 1.0 / p.scale
    for name, xyb in cases(golden).items():
        log("aq pinned", name, distance)
        _, h, w = xyb.shape
        want = oracle.fill_quant_field(xyb, p.distance, inv_scale)
        got = fe.fill_quant_field(to_dev(torch, padded(J, xyb)), w, h, p.distance,
                                  inv_scale).cpu().numpy()
        assert same_bits(got, want), (name, first_mismatch(got, want))


def _stage_inputs(J, oracle, w, h, distance, seed, fused=True, cmap=False):
    from jixel_amd.synth import photo_like
    rgb = photo_like(w, h, seed=seed)
    oracle.set_fused(fused)
    xyb = oracle.xyb_image(rgb, J.XYB_SPEC_FWD)
    p = params_for(J, distance, fused)
    qf = oracle.fill_quant_field(xyb, p.distance, 1.0 / p.scale)
    tiles = ((h + 63) // 64, (w + 63) // 64)
    if cmap:
        rng = np.random.default_rng(seed)
        ytox = rng.integers(-20, 21, size=tiles).astype(np.int8)
        ytob = rng.integers(-20, 21, size=tiles).astype(np.int8)
    else:
        ytox = np.zeros(tiles, np.int8)
        ytob = np.zeros(tiles, np.int8)
    return rgb, xyb, p, qf, ytox, ytob


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h,distance", [(64, 64, 2.35), (100, 75, 1.0), (256, 256, 2.35),
                                          (200, 136, 4.6), (300, 260, 1.6)])
def test_strategy_costs_bit_exact(J, fe, torch, oracle, w, h, distance, fused):
    log("costs", w, h, distance, fused)
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, distance, 11, fused, cmap=True)
    tables = oracle.tables_for(p)
    want = oracle.strategy_costs(xyb, qf, ytox, ytob, tables, p)
    oracle.set_fused(True)
    got = fe.strategy_costs(to_dev(torch, padded(J, xyb)), w, h, to_dev(torch, qf), p,
                            to_dev(torch, ytox), to_dev(torch, ytob))
    for g, wnt, name in zip(got, want, ["dct8", "dct16", "dct32"]):
        g = g.cpu().numpy()
        assert np.array_equal(np.isnan(g), np.isnan(wnt)), name
        m = ~np.isnan(wnt)
        assert same_bits(g[m], wnt[m]), (name, first_mismatch(g[m], wnt[m]))


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h,distance", [(64, 64, 2.35), (100, 75, 1.0), (256, 256, 2.35),
                                          (256, 200, 1.2), (2100, 72, 2.35), (130, 94, 6.0),
                                          (160, 96, 0.1), (160, 96, 0.05)])
def test_strategy_selection_bit_exact(J, fe, torch, oracle, w, h, distance, fused):
    log("select", w, h, distance, fused)
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, distance, 5, fused, cmap=True)
    tables = oracle.tables_for(p)
    cells_w, qf_w = oracle.select_strategy(xyb, qf, ytox, ytob, tables, p)
    oracle.set_fused(True)
    qf_d = to_dev(torch, qf)
    cells = fe.select_strategy(to_dev(torch, padded(J, xyb)), w, h, qf_d, p, to_dev(torch, ytox),
                               to_dev(torch, ytob)).cpu().numpy()
    assert same_bits(cells, cells_w), first_mismatch(cells, cells_w)
    assert same_bits(qf_d.cpu().numpy(), qf_w), first_mismatch(qf_d.cpu().numpy(), qf_w)
    log("  strategies", np.bincount(cells_w.reshape(-1) >> 1, minlength=5)[:5])


def high_contrast(w, h):
    """8-px and 1-px black / white checkerboards and a saturated red block:
    levels above 1024 at d = 0.1, so the rate of the decision's candidates
    takes the large-level log2 table."""
    y, x = np.mgrid[0:h, 0:w]
    img = np.zeros((h, w, 3), np.uint8)
    img[((x // 8 + y // 8) % 2) == 0] = 255
    img[:, 70:] = np.where(((x[:, 70:] + y[:, 70:]) % 2)[..., None] == 0, 255, 0)
    img[50:, :40, 0] = 255
    img[50:, :40, 1:] = 0
    return img


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("distance,content", [(0.1, "photo"), (0.05, "photo"), (0.1, "contrast"),
                                              (0.05, "contrast")])
def test_front_end_low_distance_bit_exact(J, fe, torch, oracle, distance, content, fused):
    """The finest distances: d = 0.1 still runs the decision (with the
    rerank), d = 0.05 takes the DCT8-only shortcut below 0.056713393
    (src/ac_strategy.rs:71-79); high-contrast content reaches levels above
    1024."""
    log("front_end low distance", distance, content, fused)
    from jixel_amd import synth
    w, h = 136, 104
    rgb = synth.photo_like(w, h, 4) if content == "photo" else high_contrast(w, h)
    p = J.frame_params(J.EncodeConfig(fused=fused).with_distance(distance))
    want = oracle.front_end(rgb, p, oracle.tables_for(p), threads=2)
    o = fe.run(to_dev(torch, rgb), w, h, p)
    for k in ("qf", "cells", "ytox", "ytob"):
        assert same_bits(o[k].cpu().numpy(), want[k]), (k, first_mismatch(o[k].cpu().numpy(), want[k]))
    for c in range(3):
        assert same_bits(o["levels"][c].cpu().numpy(), want["levels"][c]), c
        assert same_bits(o["coeffs"][c].cpu().numpy(), want["coeffs"][c]), c
        assert same_bits(o["dc"][c].cpu().numpy(), want["dc"][c]), c
    cells = want["cells"] >> 1
    if distance < 0.056713393:
        assert (cells == 0).all()  # every block DCT8
    else:
        assert (cells != 0).any()


def test_caller_tables_reach_every_stage(J, torch, oracle):
    """A caller's DequantMatrices (jxg_ctx_set_tables) drive selection and
    quantization exactly as the oracle with the same tables."""
    w, h = 200, 136
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, 2.35, 12, True, cmap=True)
    tables = J.default_quant_tables(p.distance).copy()
    rng = np.random.default_rng(4)
    tables *= rng.uniform(0.8, 1.25, tables.shape).astype(np.float32)
    f2 = J.FrontEnd(0, tables=tables)
    try:
        cells_w, qf_w = oracle.select_strategy(xyb, qf, ytox, ytob, tables, p)
        qf_d = to_dev(torch, qf)
        cells = f2.select_strategy(to_dev(torch, padded(J, xyb)), w, h, qf_d, p,
                                   to_dev(torch, ytox), to_dev(torch, ytob))
        assert same_bits(cells.cpu().numpy(), cells_w)
        assert same_bits(qf_d.cpu().numpy(), qf_w)
        lv_w, co_w, dc_w = oracle.forward_quantize(xyb, cells_w, qf_w, ytox, ytob, tables, p)
        o = f2.forward_quantize(to_dev(torch, padded(J, xyb)), w, h, cells, qf_d, p,
                                to_dev(torch, ytox), to_dev(torch, ytob))
        for c in range(3):
            assert same_bits(o["levels"][c].cpu().numpy(), lv_w[c]), c
            assert same_bits(o["dc"][c].cpu().numpy(), dc_w[c]), c
    finally:
        f2.close()


COV = {0: (1, 1), 1: (1, 2), 2: (2, 1), 3: (2, 2), 4: (4, 4), 8: (2, 4), 9: (4, 2), 22: (1, 4),
       23: (4, 1)}
SUB8 = [5, 6, 7, 10, 11, 12, 13, 17, 18]  # DCT4X4, 4X8, 8X4, AFV0-3, IDENTITY, DCT2X2
COV.update({s: (1, 1) for s in SUB8})
BIG = [14, 15, 16]  # DCT64X64, DCT64X32 (32 wide, 64 tall), DCT32X64 (64 wide, 32 tall)
COV.update({14: (8, 8), 15: (4, 8), 16: (8, 4)})


def random_layout(xb, yb, seed, crossing=False, sub8=False, big=False):
    """A valid strategy image: multi-block transforms inside the image and
    their 256-px group, and inside their 64-px tile unless `crossing`;
    `sub8` mixes in the single-block sub-8x8 transforms."""
    rng = np.random.default_rng(seed)
    cells = np.zeros((yb, xb), np.uint8)
    taken = np.zeros((yb, xb), bool)
    kinds = [0, 1, 2, 3, 4, 8, 9, 22, 23]
    prob = [0.3, 0.1, 0.1, 0.15, 0.1, 0.05, 0.05, 0.075, 0.075]
    if sub8:
        kinds = kinds + SUB8
        prob = [0.5 * x for x in prob] + [0.5 / len(SUB8)] * len(SUB8)
    if big:
        kinds = kinds + BIG
        prob = [0.4 * x for x in prob] + [0.2, 0.2, 0.2]
    for by in range(yb):
        for bx in range(xb):
            if taken[by, bx]:
                continue
            for _ in range(4):
                s = int(rng.choice(kinds, p=prob))
                cx, cy = COV[s]
                if bx + cx > xb or by + cy > yb:
                    continue
                if bx // 32 != (bx + cx - 1) // 32 or by // 32 != (by + cy - 1) // 32:
                    continue
                if not crossing and (bx // 8 != (bx + cx - 1) // 8 or by // 8 != (by + cy - 1) // 8):
                    continue
                if taken[by:by + cy, bx:bx + cx].any():
                    continue
                break
            else:
                s, cx, cy = 0, 1, 1
            taken[by:by + cy, bx:bx + cx] = True
            cells[by:by + cy, bx:bx + cx] = s << 1
            cells[by, bx] = (s << 1) | 1
    return cells


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h,distance,crossing", [(64, 64, 2.35, False), (130, 75, 1.0, False),
                                                    (256, 256, 2.35, True), (200, 264, 0.6, True),
                                                    (17, 9, 2.35, False)])
def test_forward_quantize_bit_exact(J, fe, torch, oracle, w, h, distance, crossing, fused):
    log("quantize", w, h, distance, crossing, fused)
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, distance, 9, fused, cmap=True)
    xb, yb = (w + 7) // 8, (h + 7) // 8
    cells = random_layout(xb, yb, seed=w * h, crossing=crossing)
    tables = oracle.tables_for(p)
    lv_w, co_w, dc_w = oracle.forward_quantize(xyb, cells, qf, ytox, ytob, tables, p)
    oracle.set_fused(True)
    o = fe.forward_quantize(to_dev(torch, padded(J, xyb)), w, h, to_dev(torch, cells),
                            to_dev(torch, qf), p, to_dev(torch, ytox), to_dev(torch, ytob))
    n_cross, bad = fe.quantize_status()
    assert bad == 0
    if not crossing:
        assert n_cross == 0
    for c in range(3):
        for name, g, wnt in [("levels", o["levels"][c], lv_w[c]), ("coeffs", o["coeffs"][c], co_w[c]),
                             ("dc", o["dc"][c], dc_w[c])]:
            g = g.cpu().numpy()
            assert same_bits(g, wnt), (name, c, first_mismatch(g, wnt))


def test_forward_quantize_composed_pins_bit_exact(J, fe, torch, oracle):
    """The single-level CfL, LLF and proportional-chroma DC cases whose
    oracle outputs follow in closed form (test_oracle_compose)."""
    from test_oracle_compose import all_inputs
    for name, a in all_inputs(J, oracle):
        log("quantize composed", name)
        lv_w, co_w, dc_w = oracle.forward_quantize(a["xyb"], a["cells"], a["qf"], a["ytox"],
                                                   a["ytob"], a["tables"], a["p"])
        oracle.set_fused(True)
        _, h, w = a["xyb"].shape
        o = fe.forward_quantize(to_dev(torch, padded(J, a["xyb"])), w, h, to_dev(torch, a["cells"]),
                                to_dev(torch, a["qf"]), a["p"], to_dev(torch, a["ytox"]),
                                to_dev(torch, a["ytob"]))
        for c in range(3):
            for what, g, wnt in [("levels", o["levels"][c], lv_w[c]),
                                 ("coeffs", o["coeffs"][c], co_w[c]), ("dc", o["dc"][c], dc_w[c])]:
                g = g.cpu().numpy()
                assert same_bits(g, wnt), (name, what, c, first_mismatch(g, wnt))


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("kind", SUB8 + ["mixed"])
def test_forward_quantize_sub8_bit_exact(J, fe, torch, oracle, kind, fused):
    """Every block one sub-8x8 transform (or a random mix with the DCTs)."""
    w, h = (200, 136) if kind == "mixed" else (72, 40)
    log("quantize sub8", kind, w, h, fused)
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, 1.4, 21, fused, cmap=True)
    xb, yb = (w + 7) // 8, (h + 7) // 8
    if kind == "mixed":
        cells = random_layout(xb, yb, seed=77, crossing=True, sub8=True, big=True)
    else:
        cells = np.full((yb, xb), (kind << 1) | 1, np.uint8)
    tables = oracle.tables_for(p)
    lv_w, co_w, dc_w = oracle.forward_quantize(xyb, cells, qf, ytox, ytob, tables, p)
    oracle.set_fused(True)
    o = fe.forward_quantize(to_dev(torch, padded(J, xyb)), w, h, to_dev(torch, cells),
                            to_dev(torch, qf), p, to_dev(torch, ytox), to_dev(torch, ytob))
    assert fe.quantize_status()[1] == 0
    for c in range(3):
        for name, g, wnt in [("levels", o["levels"][c], lv_w[c]), ("coeffs", o["coeffs"][c], co_w[c]),
                             ("dc", o["dc"][c], dc_w[c])]:
            g = g.cpu().numpy()
            assert same_bits(g, wnt), (name, c, first_mismatch(g, wnt))


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h,seed", [(64, 64, 1), (128, 72, 2), (200, 136, 3), (264, 200, 4)])
def test_forward_quantize_dct64_bit_exact(J, fe, torch, oracle, w, h, seed, fused):
    """DCT64X64 / DCT64X32 / DCT32X64 (src/dct.rs:2529-2805) beside the
    smaller transforms, on tile-aligned and crossing placements."""
    log("quantize dct64", w, h, seed, fused)
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, 2.35, 40 + seed, fused, cmap=True)
    xb, yb = (w + 7) // 8, (h + 7) // 8
    if seed == 1:
        cells = np.full((yb, xb), 14 << 1, np.uint8)
        cells[0, 0] |= 1
    else:
        cells = random_layout(xb, yb, seed=seed, crossing=seed % 2 == 0, big=True)
    assert any((c >> 1) in BIG for c in cells.reshape(-1) if c & 1)
    tables = oracle.tables_for(p)
    lv_w, co_w, dc_w = oracle.forward_quantize(xyb, cells, qf, ytox, ytob, tables, p)
    oracle.set_fused(True)
    o = fe.forward_quantize(to_dev(torch, padded(J, xyb)), w, h, to_dev(torch, cells),
                            to_dev(torch, qf), p, to_dev(torch, ytox), to_dev(torch, ytob))
    assert fe.quantize_status()[1] == 0
    for c in range(3):
        for name, g, wnt in [("levels", o["levels"][c], lv_w[c]), ("coeffs", o["coeffs"][c], co_w[c]),
                             ("dc", o["dc"][c], dc_w[c])]:
            g = g.cpu().numpy()
            assert same_bits(g, wnt), (name, c, first_mismatch(g, wnt))


def test_unsupported_cells_are_reported(J, fe, torch, oracle):
    w, h = 64, 64
    rgb, xyb, p, qf, ytox, ytob = _stage_inputs(J, oracle, w, h, 2.35, 1)
    cells = np.full((8, 8), 1, np.uint8)
    cells[0, 0] = (19 << 1) | 1  # DCT128X128: not a quantize-stage strategy
    fe.forward_quantize(to_dev(torch, padded(J, xyb)), w, h, to_dev(torch, cells), to_dev(torch, qf), p)
    assert fe.quantize_status()[1] == 1


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("name,w,h,q", [("gradient", 256, 256, 90.0), ("photo", 320, 240, 75.0),
                                        ("photo", 203, 117, 50.0), ("photo", 96, 64, 95.0)])
def test_front_end_bit_exact(J, fe, torch, oracle, name, w, h, q, fused):
    log("front_end", name, w, h, q, fused)
    from jixel_amd import synth
    rgb = (synth.gradient_sines if name == "gradient" else synth.photo_like)(w, h, 3)
    cfg = J.EncodeConfig(fused=fused).with_quality(q)
    p = J.frame_params(cfg)
    tables = oracle.tables_for(p)
    want = oracle.front_end(rgb, p, tables, threads=4)
    o = fe.run(to_dev(torch, rgb), w, h, p)
    assert same_bits(o["qf"].cpu().numpy(), want["qf"])
    assert same_bits(o["ytox"].cpu().numpy(), want["ytox"])
    assert same_bits(o["ytob"].cpu().numpy(), want["ytob"])
    assert same_bits(o["cells"].cpu().numpy(), want["cells"])
    for c in range(3):
        assert same_bits(o["levels"][c].cpu().numpy(), want["levels"][c]), c
        assert same_bits(o["coeffs"][c].cpu().numpy(), want["coeffs"][c]), c
        assert same_bits(o["dc"][c].cpu().numpy(), want["dc"][c]), c


@pytest.mark.parametrize("ch", [1, 2])
def test_front_end_gray_bit_exact(J, fe, torch, oracle, ch):
    """Gray (+ alpha) input: the achromatic path (X = 0, B = Y) end to end."""
    from jixel_amd import synth
    rgb = synth.photo_like(136, 72, 8)
    gray = rgb[:, :, :1].copy() if ch == 1 else rgb[:, :, :2].copy()
    p = J.frame_params(J.EncodeConfig().with_quality(80.0))
    want = oracle.front_end(gray, p, oracle.tables_for(p), threads=2)
    o = fe.run(to_dev(torch, gray), 136, 72, p)
    for k in ("qf", "cells", "ytox", "ytob"):
        assert same_bits(o[k].cpu().numpy(), want[k]), k
    for c in range(3):
        assert same_bits(o["levels"][c].cpu().numpy(), want["levels"][c]), c
        assert same_bits(o["dc"][c].cpu().numpy(), want["dc"][c]), c
    assert not o["levels"][0].cpu().numpy().any()  # X carries nothing


def test_host_front_end_matches_device_path(J, oracle):
    from jixel_amd import synth
    rgb = synth.photo_like(130, 70, 4)
    cfg = J.EncodeConfig().with_quality(75)
    out = J.encode_front_end(rgb, cfg)
    p = J.frame_params(cfg)
    want = oracle.front_end(rgb, p, oracle.tables_for(p), threads=2)
    assert same_bits(out.qf, want["qf"]) and same_bits(out.cells, want["cells"])
    assert same_bits(out.ytox, want["ytox"]) and same_bits(out.ytob, want["ytob"])
    for c in range(3):
        assert same_bits(out.levels[c], want["levels"][c])


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h,distance", [(64, 64, 2.35), (200, 136, 0.8), (130, 75, 1.5),
                                          (1000, 520, 2.35), (13, 70, 2.35)])
def test_cmap_bit_exact(J, fe, torch, oracle, w, h, distance, fused):
    log("cmap", w, h, distance, fused)
    rgb, xyb, p, qf, _, _ = _stage_inputs(J, oracle, w, h, distance, 31, fused)
    want_x, want_b = oracle.fill_cmap(xyb, oracle.tables_for(p), p)
    oracle.set_fused(True)
    gx, gb = fe.fill_cmap(to_dev(torch, padded(J, xyb)), w, h, p)
    assert same_bits(gx.cpu().numpy(), want_x), first_mismatch(gx.cpu().numpy(), want_x)
    assert same_bits(gb.cpu().numpy(), want_b), first_mismatch(gb.cpu().numpy(), want_b)
    assert np.any(want_x != 0) or np.any(want_b != 0) or w < 64 or h < 16


def test_error_paths(J, fe, torch):
    p = params_for(J, 2.35)
    x = torch.zeros((3, 64, 64), dtype=torch.float32, device="cuda")
    with pytest.raises(J.JxgError) as e:
        fe.fill_quant_field(x, 0, 10, 2.35, 10.0)
    assert e.value.code == -2
    with pytest.raises(J.JxgError) as e:
        fe.fill_quant_field(x, 100, 10, 2.35, 10.0)  # planes too small for 100 wide
    assert e.value.code == -1
    with pytest.raises(J.JxgError):
        J.encode_front_end(np.zeros((0, 5, 3), np.uint8))
