"""FP64 oracle of the quant-field tuning objective (no reference counterpart:
the objective is this build's own).  Pinned by central finite differences
of its own loss (h = 1e-4) and by Parseval's identity for the orthogonal
transform pair, which gives the distortion of a transform inside the image
without any inverse transform."""
import numpy as np
import pytest

LAYOUT_BLOCKS = (6, 4)  # 48 x 32 px


def layout(J):
    """DCT8 everywhere, a DCT16X16 at (0, 0), a DCT16X8 at (2, 0), a DCT8X16
    at (2, 2) and a DCT32X32-free remainder."""
    S = J.STRATEGY
    xb, yb = LAYOUT_BLOCKS
    cells = np.full((yb, xb), (S["DCT"] << 1) | 1, np.uint8)

    def put(bx, by, name, cx, cy):
        cells[by:by + cy, bx:bx + cx] = S[name] << 1
        cells[by, bx] |= 1
    put(0, 0, "DCT16X16", 2, 2)
    put(2, 0, "DCT16X8", 1, 2)   # 8 wide, 16 tall
    put(2, 2, "DCT8X16", 2, 1)   # 16 wide, 8 tall
    put(4, 2, "DCT16X16", 2, 2)
    return cells


def setup(J, oracle, w, h, seed=3):
    from jixel_amd import synth
    rgb = synth.photo_like(w, h, seed)
    xyb = oracle.xyb_image(rgb, J.XYB_SPEC_FWD)
    p = J.frame_params(J.EncodeConfig().with_quality(75))
    return xyb, p, oracle.tables_for(p)


def theta_for(cells, rng):
    th = rng.uniform(20.0, 80.0, cells.shape)
    # covered blocks share their transform's value, as after the front end
    S_cov = {}
    yb, xb = cells.shape
    for by in range(yb):
        for bx in range(xb):
            if cells[by, bx] & 1:
                S_cov[(bx, by)] = th[by, bx]
    return th


@pytest.mark.parametrize("lam", [0.0, 0.01])
def test_adjoint_matches_central_differences(oracle, lam):
    import jixel_amd as J
    cells = layout(J)
    xyb, p, tables = setup(J, oracle, 45, 29)  # ragged: transforms reach past the image
    rng = np.random.default_rng(5)
    th = theta_for(cells, rng)
    loss, g = oracle.tune_loss_grad(xyb, cells, th, tables, p, lam, 7, 2)
    assert loss > 0
    hstep = 1e-4
    for by in range(cells.shape[0]):
        for bx in range(cells.shape[1]):
            tp, tm = th.copy(), th.copy()
            tp[by, bx] += hstep
            tm[by, bx] -= hstep
            lp, _ = oracle.tune_loss_grad(xyb, cells, tp, tables, p, lam, 7, 2, with_grad=False)
            lm, _ = oracle.tune_loss_grad(xyb, cells, tm, tables, p, lam, 7, 2, with_grad=False)
            fd = (lp - lm) / (2 * hstep)
            assert abs(fd - g[by, bx]) <= 1e-6 * max(1.0, abs(loss)) + 1e-5 * abs(fd), (bx, by)


def test_distortion_obeys_parseval(oracle):
    """With lambda = 0 and every transform inside the image, the loss is
    sum_c w_c * (w h) * sum_k (s_k u_k)^2 over the noisy coefficients."""
    import jixel_amd as J
    cells = layout(J)
    xyb, p, tables = setup(J, oracle, 48, 32)
    th = np.full(cells.shape, 37.0)
    loss, _ = oracle.tune_loss_grad(xyb, cells, th, tables, p, 0.0, 11, 0)
    S = J.STRATEGY
    cov = {S["DCT"]: (1, 1), S["DCT16X16"]: (2, 2), S["DCT16X8"]: (1, 2), S["DCT8X16"]: (2, 1)}
    kinds = {S["DCT"]: 0, S["DCT16X8"]: 1, S["DCT8X16"]: 1, S["DCT16X16"]: 2}
    qm = [1.25 ** (p.x_qm_scale - 2.0), 1.0, None]
    want = 0.0
    xb = cells.shape[1]
    for by in range(cells.shape[0]):
        for bx in range(xb):
            if not cells[by, bx] & 1:
                continue
            s = int(cells[by, bx] >> 1)
            cx, cy = cov[s]
            n = 64 * cx * cy
            kind = kinds[s]
            off, size = J.qt_offset(kind), J.qt_size(kind)
            for c in range(3):
                inv = tables[off + (3 + c) * size: off + (4 + c) * size].astype(np.float64)
                qmc = qm[c] if c != 2 else float(np.float32(1.0))
                if c == 2:
                    # b_qm_mul at b_qm_scale 2 is 1 (the default table index)
                    assert p.b_qm_scale == 2
                qs = p.scale * 37.0 * qmc
                e2 = 0.0
                for k in range(n):
                    v, u = storage_to_spatial(s, S, cx * 8, cy * 8, k)
                    if v < cy and u < cx:
                        continue
                    st = 1.0 / (inv[k] * qs)
                    un = oracle.tune_noise(11, 0, c, by * xb + bx, k)
                    e2 += (st * un) ** 2
                want += p.channel_weight[c] * (cx * 8) * (cy * 8) * e2
    assert abs(loss - want) <= 1e-9 * want


def storage_to_spatial(s, S, w, h, k):
    if s == S["DCT"]:
        return k & 7, k >> 3
    if s in (S["DCT8X16"], S["DCT16X32"], S["DCT8X32"]):
        return k // w, k % w
    return k % h, k // h


def test_noise_is_uniform_and_deterministic(oracle):
    u = np.array([oracle.tune_noise(3, 1, c, b, k) for c in range(3) for b in range(20)
                  for k in range(64)])
    assert u.min() >= -0.5 and u.max() < 0.5
    assert abs(u.mean()) < 0.02 and abs(u.var() - 1.0 / 12.0) < 0.005
    assert oracle.tune_noise(3, 1, 2, 7, 5) == oracle.tune_noise(3, 1, 2, 7, 5)
    assert oracle.tune_noise(3, 1, 2, 7, 5) != oracle.tune_noise(3, 2, 2, 7, 5)
