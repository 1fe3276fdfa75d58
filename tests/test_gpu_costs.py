"""GPU parity of the per-site candidate costs (every strategy up to 64x64,
sub-8x8 and AFV included) against the oracle's strategy_cost, bit for bit,
and against the strategy-search kernel's own DCT8 costs."""
import numpy as np
import pytest

from test_gpu_parity import first_mismatch, same_bits, to_dev

pytestmark = pytest.mark.gpu


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


ALL = ["DCT", "DCT16X8", "DCT8X16", "DCT16X16", "DCT32X32", "DCT32X16", "DCT16X32", "DCT32X8",
       "DCT8X32", "DCT64X64", "DCT64X32", "DCT32X64", "IDENTITY", "DCT2X2", "DCT4X4", "DCT4X8",
       "DCT8X4", "AFV0", "AFV1", "AFV2", "AFV3"]

COV = {"DCT": (1, 1), "DCT16X8": (1, 2), "DCT8X16": (2, 1), "DCT16X16": (2, 2),
       "DCT32X32": (4, 4), "DCT32X16": (2, 4), "DCT16X32": (4, 2), "DCT32X8": (1, 4),
       "DCT8X32": (4, 1), "DCT64X64": (8, 8), "DCT64X32": (4, 8), "DCT32X64": (8, 4)}


def f32_fmla(a, b, c, fused):
    a, b, c = np.float32(a), np.float32(b), np.float32(c)
    if fused:
        return np.float32(np.float64(a) * np.float64(b) + np.float64(c))
    return np.float32(a * b + c)


def sites_for(name, w, h, rng, n=24):
    cx, cy = COV.get(name, (1, 1))
    xb, yb = (w + 7) // 8, (h + 7) // 8
    s, ph = ((w + 63) // 64 * 64, (h + 63) // 64 * 64)
    out = []
    for _ in range(n):
        bx = int(rng.integers(0, max(1, xb - cx + 1))) // cx * cx
        by = int(rng.integers(0, max(1, yb - cy + 1))) // cy * cy
        out.append((bx, by))
    # one site reaching past the image (visible fraction < 1), inside the planes
    bx, by = (xb - 1) // cx * cx, (yb - 1) // cy * cy
    if (bx + cx) * 8 <= s and (by + cy) * 8 <= ph:
        out.append((bx, by))
    return out


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("name", ALL)
def test_site_costs_bit_exact(J, fe, torch, oracle, name, fused):
    from jixel_amd import STRATEGY, synth
    w, h = 200, 136
    print("site_cost", name, fused, flush=True)
    rgb = synth.photo_like(w, h, 21)
    d = 2.35 if name not in ("IDENTITY", "DCT2X2") else 1.2
    p = J.frame_params(J.EncodeConfig(fused=fused).with_distance(d))
    xyb_d = fe.xyb(to_dev(torch, rgb), w, h)
    xyb = xyb_d[:, :h, :w].cpu().numpy()
    rng = np.random.default_rng(len(name) * 7 + fused)
    tiles = ((h + 63) // 64, (w + 63) // 64)
    ytox = rng.integers(-40, 41, tiles).astype(np.int8)
    ytob = rng.integers(-40, 41, tiles).astype(np.int8)
    sites = sites_for(name, w, h, rng)
    qac = (p.scale * rng.integers(4, 120, len(sites))).astype(np.float32)
    tables = oracle.tables_for(p)
    s = STRATEGY[name]
    for meta_r in (4.0, 0.0):
        got = fe.strategy_cost_sites(xyb_d, w, h, p, s, sites, qac, meta_r=meta_r,
                                     ytox=to_dev(torch, ytox), ytob=to_dev(torch, ytob)).cpu().numpy()
        oracle.set_fused(fused)
        want = []
        for (bx, by), q in zip(sites, qac):
            t = ytox[min(by // 8, tiles[0] - 1), min(bx // 8, tiles[1] - 1)]
            u = ytob[min(by // 8, tiles[0] - 1), min(bx // 8, tiles[1] - 1)]
            inv_cf = np.float32(1.0) / np.float32(p.cfl_color_factor)
            cm = [f32_fmla(t, inv_cf, p.cfl_base_x, fused), 0.0,
                  f32_fmla(u, inv_cf, p.cfl_base_b, fused)]
            want.append(oracle.strategy_cost(xyb, s, bx * 8, by * 8, float(q), cm, tables, p,
                                             meta_r=meta_r))
        oracle.set_fused(True)
        want = np.asarray(want, np.float32)
        assert same_bits(got, want), (meta_r, first_mismatch(got, want))


def test_site_costs_match_square_search(J, fe, torch, oracle):
    """The square costs of the search kernel (jxg_strategy_costs_dev) and the
    site kernel agree at the search's own sites and quantizers."""
    from jixel_amd import STRATEGY, synth
    w, h = 256, 192
    rgb = synth.photo_like(w, h, 5)
    p = J.frame_params(J.EncodeConfig().with_quality(75))
    xyb = fe.xyb(to_dev(torch, rgb), w, h)
    qf = fe.fill_quant_field(xyb, w, h, p.distance, 1.0 / p.scale)
    c8, c16, c32 = fe.strategy_costs(xyb, w, h, qf, p)
    c8, c16, c32 = c8.cpu().numpy(), c16.cpu().numpy(), c32.cpu().numpy()
    q = qf.cpu().numpy()
    xb, yb = (w + 7) // 8, (h + 7) // 8
    sites = [(bx, by) for by in range(yb) for bx in range(xb)]
    qac = np.float32(p.scale) * q[[b for _, b in sites], [a for a, _ in sites]].astype(np.float32)
    got = fe.strategy_cost_sites(xyb, w, h, p, STRATEGY["DCT"], sites, qac).cpu().numpy()
    want = c8.reshape(-1)
    ok = np.isfinite(want)  # blocks the search evaluated
    assert ok.sum() > 0.9 * want.size
    assert same_bits(got[ok], want[ok]), first_mismatch(got[ok], want[ok])
