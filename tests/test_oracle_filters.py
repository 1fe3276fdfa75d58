"""Oracle pins for the whole-plane filters (gaborish pre-inversion).

The gaborish restatement is checked against what the mathematics fixes: an
independent FP64 5x5 convolution with an explicit weight matrix and mirror
reflection, the unit DC gain of the normalised kernel, and the reference's
own constant-input test (src/gaborish.rs:327)."""
import numpy as np
import pytest

K_GABORISH = [-0.09495815671340026, -0.041031725066768575, 0.013710004822696948,
              0.006510206083837737, -0.0014789063378272242]  # src/gaborish.rs:32-38


def kernel5_f64(mul):
    """Explicit 5x5 weight matrix (offsets -2..2) of the normalised kernel."""
    k0, k1, k2, k3, k4 = [mul * k for k in K_GABORISH]
    m = np.zeros((5, 5))
    m[2, 2] = 1.0
    for dy, dx, wgt in [(0, 1, k0), (1, 1, k1), (0, 2, k2), (2, 2, k4), (1, 2, k3), (2, 1, k3)]:
        for sy in (-1, 1):
            for sx in (-1, 1):
                for a, b in ((dy, dx), (dx, dy)):
                    m[2 + sy * a, 2 + sx * b] = wgt
    return m / m.sum()


def reflect(i, n):
    while i < 0 or i >= n:
        i = -i - 1 if i < 0 else 2 * n - i - 1
    return i


def conv_f64(plane, mul):
    h, w = plane.shape
    k = kernel5_f64(mul)
    p = plane.astype(np.float64)
    out = np.zeros_like(p)
    ry = np.array([[reflect(y + d, h) for d in range(-2, 3)] for y in range(h)])
    rx = np.array([[reflect(x + d, w) for d in range(-2, 3)] for x in range(w)])
    for a in range(5):
        for b in range(5):
            out += k[a, b] * p[ry[:, a]][:, rx[:, b]]
    return out


def test_kernel_matrix_has_the_reference_taps():
    m = kernel5_f64(1.0)
    assert np.count_nonzero(m) == 25  # every 5x5 position carries a tap
    assert np.allclose(m, m.T) and np.allclose(m, m[::-1]) and np.allclose(m, m[:, ::-1])


@pytest.mark.parametrize("mul", [0.9908511, 0.0, 0.5, 2.0])
def test_weights_match_fp64_and_sum_to_one(oracle, mul):
    w = oracle.gaborish_weights(mul).astype(np.float64)
    k = kernel5_f64(mul)
    want = [k[2, 2], k[2, 3], k[3, 3], k[2, 4], k[3, 4], k[4, 4]]  # c, r, d, R, L, D
    assert np.allclose(w, want, rtol=2e-6, atol=1e-9)
    assert abs(w[0] + 4 * w[1] + 4 * w[2] + 4 * w[3] + 8 * w[4] + 4 * w[5] - 1.0) < 1e-6


@pytest.mark.parametrize("w,h", [(1, 1), (2, 3), (5, 4), (16, 17), (33, 7), (64, 64)])
def test_gaborish_matches_fp64_convolution(oracle, w, h):
    rng = np.random.default_rng(w * 100 + h)
    xyb = rng.uniform(-0.1, 0.9, size=(3, h, w)).astype(np.float32)
    for fused in (True, False):
        oracle.set_fused(fused)
        got = oracle.gaborish_inverse(xyb)
        for c in range(3):
            want = conv_f64(xyb[c], 0.9908511)
            assert np.max(np.abs(got[c] - want)) < 4e-6, (w, h, c, fused)
    oracle.set_fused(True)


def test_gaborish_keeps_constant_planes(oracle):
    """src/gaborish.rs:327: a constant 16x16 image is unchanged within 1e-5."""
    xyb = np.full((3, 16, 16), 0.5, np.float32)
    got = oracle.gaborish_inverse(xyb)
    assert np.max(np.abs(got - 0.5)) < 1e-5


def test_gaborish_at_zero_mul_is_identity(oracle):
    rng = np.random.default_rng(3)
    xyb = rng.uniform(0, 1, size=(3, 9, 11)).astype(np.float32)
    assert np.array_equal(oracle.gaborish_inverse(xyb, 0.0), xyb)
