"""Quant-field tuning on the GPU: the FP32 loss and hand-written adjoint
against the FP64 oracle (loss and gradient) and against FP64 central
finite differences (h = 1e-4) on 64 random blocks; Adam's integer field."""
import numpy as np
import pytest

from test_gpu_parity import to_dev

# covered_blocks_x_of / _y_of (src/dc_group_data.rs:152-169)
COV_X = [1, 1, 2, 2, 4, 1, 1, 1, 2, 4, 1, 1, 1, 1, 8, 4, 8, 1, 1, 16, 8, 16, 1, 4]
COV_Y = [1, 2, 1, 2, 4, 1, 1, 1, 4, 2, 1, 1, 1, 1, 8, 8, 4, 1, 1, 16, 16, 8, 4, 1]

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


def front(J, fe, torch, w, h, q, seed):
    from jixel_amd import synth
    rgb = synth.photo_like(w, h, seed)
    p = J.frame_params(J.EncodeConfig().with_quality(q))
    d_rgb = to_dev(torch, rgb)
    o = fe.run(d_rgb, w, h, p)
    xyb = fe.xyb(d_rgb, w, h)
    return xyb, o["cells"], o["qf"], p


@pytest.mark.parametrize("w,h,q,lam", [(256, 192, 75.0, 0.01), (203, 117, 60.0, 0.01),
                                       (128, 96, 90.0, 0.0)])
def test_loss_and_gradient_match_fp64(J, fe, torch, oracle, w, h, q, lam):
    print("tune", w, h, q, lam, flush=True)
    xyb, cells, qf, p = front(J, fe, torch, w, h, q, 5)
    tuner = J.QfTuner(fe, xyb, w, h, cells, p, lam=lam, seed=9)
    try:
        # transforms crossing a 32x32 region (unaligned re-tiled pieces) are
        # left out, in the kernels and in the oracle alike
        cn = cells.cpu().numpy()
        crossing = 0
        for by, bx in zip(*np.nonzero(cn & 1)):
            s = int(cn[by, bx] >> 1)
            crossing += int((bx & 3) + COV_X[s] > 4 or (by & 3) + COV_Y[s] > 4)
        assert tuner.excluded() == crossing
        rng = np.random.default_rng(1)
        theta = qf.float() * torch.tensor(rng.uniform(0.8, 1.25, qf.shape).astype(np.float32),
                                          device=qf.device)
        loss, grad = tuner.loss_grad(theta, step=3)
        x = xyb[:, :h, :w].cpu().numpy()
        c = cells.cpu().numpy()
        th = theta.cpu().numpy().astype(np.float64)
        want_loss, want_grad = oracle.tune_loss_grad(x, c, th, oracle.tables_for(p), p, lam,
                                                     9, 3)
        g = grad.cpu().numpy().astype(np.float64)
        assert abs(loss - want_loss) <= 2e-5 * abs(want_loss), (loss, want_loss)
        scale = np.abs(want_grad).max()
        err = np.abs(g - want_grad).max()
        print("loss", loss, want_loss, "grad scale", scale, "max err", err, flush=True)
        assert err <= 2e-3 * scale
        # FP64 central differences on 64 random blocks (the spec's check)
        picks = rng.choice(th.size, 64, replace=False)
        hstep = 1e-4
        for i in picks:
            by, bx = divmod(int(i), th.shape[1])
            tp, tm = th.copy(), th.copy()
            tp[by, bx] += hstep
            tm[by, bx] -= hstep
            lp, _ = oracle.tune_loss_grad(x, c, tp, oracle.tables_for(p), p, lam, 9, 3, False)
            lm, _ = oracle.tune_loss_grad(x, c, tm, oracle.tables_for(p), p, lam, 9, 3, False)
            fd = (lp - lm) / (2 * hstep)
            assert abs(g[by, bx] - fd) <= 2e-3 * scale + 1e-3 * abs(fd), (bx, by, g[by, bx], fd)
    finally:
        tuner.close()


def test_adam_gives_a_valid_lower_loss_field(J, fe, torch):
    w, h = 320, 256
    xyb, cells, qf, p = front(J, fe, torch, w, h, 75.0, 7)
    tuner = J.QfTuner(fe, xyb, w, h, cells, p, lam=0.01, seed=3)
    try:
        out, theta, trace = tuner.run(qf, steps=20, lr=0.05)
        print("loss trace", trace[0], trace[-1], flush=True)
        assert trace[-1] < trace[0]
        q = out.cpu().numpy()
        assert q.min() >= 1 and q.max() <= 255
        c = cells.cpu().numpy()
        yb, xb = c.shape
        # one value per transform
        for by in range(yb):
            for bx in range(xb):
                if c[by, bx] & 1:
                    s = int(c[by, bx] >> 1)
                    nx, ny = COV_X[s], COV_Y[s]
                    assert (q[by:by + ny, bx:bx + nx] == q[by, bx]).all()
        # the tuned field quantizes like any other
        o = fe.forward_quantize(xyb, w, h, cells, out, p)
        torch.cuda.synchronize()
        assert fe.quantize_status()[1] == 0
    finally:
        tuner.close()
