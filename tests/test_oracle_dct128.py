"""Oracle of the 128 family's forward transforms (src/dct128.rs:330-468):
its butterfly constants equal the reference's WC128 literals, and all three
shapes agree with the FP64 closed form within 1e-5 of the largest
coefficient, in the reference's storage order."""
import math

import numpy as np
import pytest


def cmat(n):
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    m = np.cos(np.pi * k * (2 * j + 1) / (2 * n))
    m[1:] *= math.sqrt(2.0)
    return m


def test_wc128_equals_reference_literals(oracle, golden):
    assert np.array_equal(oracle.wc128(), np.asarray(golden["WC128"], np.float32))


@pytest.mark.parametrize("kind,rows,cols", [(0, 128, 128), (1, 128, 64), (2, 64, 128)])
def test_dct128_matches_fp64(oracle, kind, rows, cols):
    rng = np.random.default_rng(kind)
    x = rng.standard_normal((rows, cols)).astype(np.float32)
    y = cmat(rows) @ x.astype(np.float64) @ cmat(cols).T / (rows * cols)  # [v, u]
    got = oracle.dct128(kind, x)
    # storage: 128x128 and 128 tall x 64 wide are u-major (u * 128 + v),
    # 64 tall x 128 wide is v-major (v * 128 + u)
    want = y.T.reshape(-1) if kind in (0, 1) else y.reshape(-1)
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


def test_dct128_of_a_constant_is_its_dc(oracle):
    x = np.full((128, 128), 0.375, np.float32)
    got = oracle.dct128(0, x)
    assert abs(got[0] - 0.375) < 1e-7
    assert np.abs(got[1:]).max() < 1e-7
