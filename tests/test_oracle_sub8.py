"""Oracle pins for the sub-8x8 transforms (IDENTITY, DCT2X2, DCT4X4,
DCT4X8, DCT8X4, AFV0-3; src/dct.rs:185-330,1061-1340, src/afv.rs:207-256).

Each transform is checked against an FP64 closed form built from the
textbook definitions (the scaled DCT-II cosine matrix, Haar pyramids, the
AFV basis held by the reference) and the storage layout the reference
documents, not against a retyped copy of the oracle's loops."""
import ctypes
import math

import numpy as np
import pytest

S = dict(DCT4X4=5, DCT4X8=6, DCT8X4=7, AFV0=10, AFV1=11, AFV2=12, AFV3=13, IDENTITY=17,
         DCT2X2=18)


def C(n):
    """Scaled DCT-II: X[k] = c_k sum x[i] cos(pi k (2i+1) / 2n), c_0 = 1, c_k = sqrt 2."""
    k = np.arange(n)[:, None]
    i = np.arange(n)[None, :]
    m = np.cos(np.pi * k * (2 * i + 1) / (2 * n))
    m[1:] *= math.sqrt(2.0)
    return m


def dct2(block):
    h, w = block.shape
    return C(h) @ block @ C(w).T / (w * h)  # [vf, hf]


def block(seed):
    return np.random.default_rng(seed).uniform(-1, 1, (8, 8)).astype(np.float32)


def fwd(oracle, name, px):
    return oracle.forward_transform(S[name], px, 0, 0).astype(np.float64)


def close(got, want, tol=2e-6):
    return np.max(np.abs(got - want)) <= tol


def test_dct4x4_closed_form(oracle):
    px = block(1)
    got = fwd(oracle, "DCT4X4", px).reshape(8, 8)
    p = px.astype(np.float64)
    want = np.zeros((8, 8))
    dcs = {}
    for qy in range(2):
        for qx in range(2):
            d = dct2(p[4 * qy:4 * qy + 4, 4 * qx:4 * qx + 4])
            for vf in range(4):
                for hf in range(4):
                    want[qy + 2 * hf, qx + 2 * vf] = d[vf, hf]
            dcs[qy, qx] = d[0, 0]
    b00, b01, b10, b11 = dcs[0, 0], dcs[0, 1], dcs[1, 0], dcs[1, 1]
    want[0, 0] = (b00 + b01 + b10 + b11) / 4
    want[0, 1] = (b00 + b01 - b10 - b11) / 4
    want[1, 0] = (b00 - b01 + b10 - b11) / 4
    want[1, 1] = (b00 - b01 - b10 + b11) / 4
    assert close(got, want)
    assert abs(got[0, 0] - p.mean()) < 1e-6


def test_dct4x8_closed_form(oracle):
    """Two 4-tall x 8-wide halves, interleaved on even / odd rows."""
    px = block(2)
    got = fwd(oracle, "DCT4X8", px).reshape(8, 8)
    p = px.astype(np.float64)
    want = np.zeros((8, 8))
    d = [dct2(p[4 * y:4 * y + 4, :]) for y in range(2)]
    for y in range(2):
        for vf in range(4):
            want[y + 2 * vf, :] = d[y][vf, :]
    want[0, 0] = (d[0][0, 0] + d[1][0, 0]) / 2
    want[1, 0] = (d[0][0, 0] - d[1][0, 0]) / 2
    assert close(got, want)


def test_dct8x4_closed_form(oracle):
    """Two 8-tall x 4-wide halves; half x's (vf, hf) lands at row x + 2 hf,
    column vf."""
    px = block(3)
    got = fwd(oracle, "DCT8X4", px).reshape(8, 8)
    p = px.astype(np.float64)
    want = np.zeros((8, 8))
    d = [dct2(p[:, 4 * x:4 * x + 4]) for x in range(2)]
    for x in range(2):
        for hf in range(4):
            want[x + 2 * hf, :] = d[x][:, hf]
    want[0, 0] = (d[0][0, 0] + d[1][0, 0]) / 2
    want[1, 0] = (d[0][0, 0] - d[1][0, 0]) / 2
    assert close(got, want)


def haar_level(a):
    """One 2x2 Haar level of an S x S array -> (avg, horiz, vert, diag) quadrants."""
    c00, c01, c10, c11 = a[0::2, 0::2], a[0::2, 1::2], a[1::2, 0::2], a[1::2, 1::2]
    return ((c00 + c01 + c10 + c11) / 4, (c00 + c01 - c10 - c11) / 4,
            (c00 - c01 + c10 - c11) / 4, (c00 - c01 - c10 + c11) / 4)


def test_dct2x2_is_a_three_level_haar_pyramid(oracle):
    px = block(4)
    got = fwd(oracle, "DCT2X2", px).reshape(8, 8)
    cur = px.astype(np.float64)
    want = np.zeros((8, 8))
    s = 8
    while s >= 2:
        a, b, c, d = haar_level(cur)
        h = s // 2
        want[:h, :h], want[:h, h:s], want[h:s, :h], want[h:s, h:s] = a, b, c, d
        cur = a
        s = h
    assert close(got, want)


def inv_identity(co):
    """Pixels from IDENTITY coefficients: quadrant means from the 2x2
    Hadamard in {0, 1, 8, 9}; residual of quadrant pixel (iy, ix) at
    (qy + 2 iy, qx + 2 ix), the (0, 0) residual parked at (qy + 2, qx + 2),
    the (1, 1) pixel is the anchor."""
    g = co.reshape(8, 8)
    means = [g[0, 0] + g[0, 1] + g[1, 0] + g[1, 1], g[0, 0] + g[0, 1] - g[1, 0] - g[1, 1],
             g[0, 0] - g[0, 1] + g[1, 0] - g[1, 1], g[0, 0] - g[0, 1] - g[1, 0] + g[1, 1]]
    out = np.zeros((8, 8))
    for qy in range(2):
        for qx in range(2):
            res = np.zeros((4, 4))
            for iy in range(4):
                for ix in range(4):
                    if (iy, ix) == (1, 1):
                        continue
                    res[iy, ix] = g[qy + 2, qx + 2] if (iy, ix) == (0, 0) else g[qy + 2 * iy, qx + 2 * ix]
            anchor = means[qy * 2 + qx] - res.sum() / 16
            out[4 * qy:4 * qy + 4, 4 * qx:4 * qx + 4] = anchor + res
    return out


def test_identity_round_trip_and_mean(oracle):
    px = block(5)
    got = fwd(oracle, "IDENTITY", px)
    assert close(inv_identity(got), px.astype(np.float64), 4e-6)
    assert abs(got[0] - px.astype(np.float64).mean()) < 1e-6


def test_afv_table_is_the_reference_basis(oracle, golden):
    L = oracle.lib()
    L.jxo_afv_basis.restype = ctypes.POINTER(ctypes.c_float)
    got = np.ctypeslib.as_array(L.jxo_afv_basis(), shape=(256,)).reshape(16, 16)
    want = np.array(golden["AFV_BASIS_TRANSPOSE"], np.float64).astype(np.float32)
    assert np.array_equal(got, want)
    b = got.astype(np.float64)
    assert np.abs(b.T @ b - np.eye(16)).max() < 1e-6  # orthonormal
    assert np.all(b[:, 0] == 0.25)


@pytest.mark.parametrize("kind", range(4))
def test_afv_closed_form(oracle, golden, kind):
    """AFV quadrant (mirrored so the block corner maps to basis pixel 0) on
    even rows / even columns, the adjacent quadrant's DCT4x4 on even rows /
    odd columns, the other half's DCT4x8 on odd rows; coefficient 0 becomes
    the block mean."""
    px = block(10 + kind)
    p = px.astype(np.float64)
    got = fwd(oracle, f"AFV{kind}", px).reshape(8, 8)
    B = np.array(golden["AFV_BASIS_TRANSPOSE"], np.float64)
    ax, ay = kind & 1, kind >> 1
    quad = p[4 * ay:4 * ay + 4, 4 * ax:4 * ax + 4]
    if ay:
        quad = quad[::-1, :]
    if ax:
        quad = quad[:, ::-1]
    coeff = quad.reshape(16) @ B
    want = np.zeros((8, 8))
    for iy in range(4):
        for ix in range(4):
            want[2 * iy, 2 * ix] = coeff[iy * 4 + ix]
    qx = 0 if ax else 4
    d44 = dct2(p[4 * ay:4 * ay + 4, qx:qx + 4])
    for vf in range(4):
        for hf in range(4):
            want[2 * hf, 2 * vf + 1] = d44[vf, hf]
    hy = 0 if ay else 4
    d48 = dct2(p[hy:hy + 4, :])
    for vf in range(4):
        want[2 * vf + 1, :] = d48[vf, :]
    b00, b01, b10 = want[0, 0] / 4, want[0, 1], want[1, 0]
    want[0, 0] = (b00 + b01 + 2 * b10) / 4
    want[0, 1] = (b00 - b01) / 2
    want[1, 0] = (b00 + b01 - 2 * b10) / 4
    assert close(got, want, 4e-6)
    assert abs(got[0, 0] - p.mean()) < 1e-6


@pytest.mark.parametrize("name", sorted(S))
def test_dc_of_constant_block(oracle, name):
    px = np.full((8, 8), 0.625, np.float32)
    got = fwd(oracle, name, px)
    assert abs(got[0] - 0.625) < 1e-6
    assert np.max(np.abs(got[1:])) < 1e-6 if name != "IDENTITY" else True
