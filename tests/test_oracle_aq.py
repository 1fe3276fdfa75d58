"""Pins of the oracle's adaptive quantization field below d = 1.75, where the
masking is on (src/adaptive_quant.rs:547-729).

The check is an FP64 model of the same chain, written as array operations
rather than pixel loops: the four-row accumulator and the 4x subsampling are
one reshape-sum, FuzzyErosion's sorting network and min-insertion are a sort
of the nine neighbours, the per-block modulations are reshape-sums over
8x8 blocks, and fast_exp2 / dirty_log2f are exact powers and logarithms.
Every constant comes from the reference's source text (tests/golden,
written by make_reference_tables.py), not from the oracle.

The images are chosen so that each term is isolated: constant planes (the
pre-map, erosion and block sums reduce to counts, checked against a scalar
closed form as well), luma steps across and along rows inside a block (the
neighbour differences, the valmin_y clamp of the HF term, the 0.2 clamp of
the squared difference on the dark side of a step, the erosion's four
smallest of nine), blue planes above luma by a fraction of k_limit, by more
than k_limit (the all-blue fold) and below it, and photo content.  A dropped
term, a wrong neighbour, a transposed difference, a sum taken as a mean or
a mistyped constant moves the field by far more than the tolerance."""
import math

import numpy as np
import pytest

DISTANCES = (0.5, 1.0, 1.5)


class Model:
    def __init__(self, golden):
        c = golden["AQ"]
        self.c = c
        g = lambda k, i=0: c[k][i]
        self.k_sg_mul = g("K_SG_MUL")
        self.k_sg_mul2 = c["K_SG_MUL2"][0] / c["K_SG_MUL2"][1]
        self.k_log2 = g("K_LOG2")
        self.k_sg_ret_mul = self.k_sg_mul2 * g("K_SG_RET_MUL") * self.k_log2
        self.k_sg_v_offset = g("K_SG_V_OFFSET")
        self.k_ac_quant = g("K_AC_QUANT")
        self.gamma_offset = g("MATCH_GAMMA_OFFSET")
        self.eps = g("ratio_cubic_to_simple_gamma.k_epsilon")
        self.num3 = g("ratio_cubic_to_simple_gamma.k_num_mul")

    def ratio(self, v, invert):
        """ratio_cubic_to_simple_gamma (src/adaptive_quant.rs:52-63)."""
        v = np.maximum(v, 0.0)
        k_num_mul = self.k_sg_ret_mul * self.num3 * self.k_sg_mul
        k_v_offset = self.k_sg_v_offset * self.k_log2 + self.eps
        k_den_mul = self.k_log2 * self.k_sg_mul
        num = k_num_mul * v * v + self.eps
        den = k_den_mul * v * v * v + k_v_offset
        return num / den if invert else den / num

    def masking_sqrt(self, v):
        c = self.c
        mul_v = c["masking_sqrt.k_mul"][0] * c["masking_sqrt.mul_v"][0]
        return 0.25 * np.sqrt(v * math.sqrt(mul_v) + c["masking_sqrt.k_log_offset"][0])

    def compute_mask(self, out_val):
        c = self.c
        g = lambda k: c["compute_mask." + k][0]
        v1 = np.maximum(out_val * g("k_mul0"), g("v1"))
        v2 = 1.0 / (v1 + g("k_offset2"))
        v3 = 1.0 / (v1 * v1 + g("k_offset3"))
        v4 = 1.0 / (v1 * v1 + g("k_offset4") * g("k_offset3"))
        return g("k_mul4") * v4 + g("k_mul2") * v2 + g("k_mul3") * v3 + g("k_base")

    def schedule(self, name, d):
        knots = self.c["AQ_KNOTS"]
        return float(np.interp(d, knots, self.c[name]))

    def erosion_weights(self, d):
        c = self.c
        fe_mul = (2.0 - d) * 0.5 if d < 2.0 else 0.0
        w = np.array(c["fill_quant_field_scalar.fe_base"]) + fe_mul * np.array(
            c["fill_quant_field_scalar.fe_add"])
        return w * (c["fill_quant_field_scalar.k_total"][0] / w.sum())

    def pre_map(self, Y):
        H, W = Y.shape
        P = np.pad(Y, 1, mode="edge")
        base = 0.25 * (P[2:, 1:-1] + P[:-2, 1:-1] + P[1:-1, :-2] + P[1:-1, 2:])
        diff = (self.ratio(Y + self.gamma_offset, False) * (Y - base)) ** 2
        diff = np.minimum(diff, self.c["fill_quant_field_scalar.diff_clamp"][0])
        ms = self.masking_sqrt(diff)
        return ms.reshape(H // 4, 4, W // 4, 4).sum(axis=(1, 3)) * 0.25

    def aq_map(self, pre, d):
        ph, pw = pre.shape
        P = np.pad(pre, 1, mode="edge")
        nb = np.stack([P[1 + dy:1 + dy + ph, 1 + dx:1 + dx + pw]
                       for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
        four = np.sort(nb, axis=0)[:4]
        v = np.tensordot(self.erosion_weights(d), four, axes=1)
        return v.reshape(ph // 2, 2, pw // 2, 2).sum(axis=(1, 3))

    @staticmethod
    def block_sum(a):
        H, W = a.shape
        return a.reshape(H // 8, 8, W // 8, 8).sum(axis=(1, 3))

    def gamma_term(self, X, Y):
        iny = Y + self.c["gamma_modulation.k_bias"][0]
        r = 0.5 * (self.ratio(iny - X, True) + self.ratio(iny + X, True))
        return self.c["gamma_modulation.k_gam"][0] * np.log2(self.block_sum(r) / 64.0)

    def hf_sum(self, Y):
        H, W = Y.shape
        vmin = self.c["hf_modulation.valmin_y"][0]
        hx = np.minimum(vmin, np.abs(Y[:, :-1] - Y[:, 1:]))
        hx[:, 7::8] = 0.0  # the pair (7, 8) crosses into the next block
        hy = np.minimum(vmin, np.abs(Y[:-1, :] - Y[1:, :]))
        hy[7::8, :] = 0.0  # the last row is compared with itself
        h = np.zeros_like(Y)
        h[:, :-1] += hx
        h[:-1, :] += hy
        return self.block_sum(h)

    def blue_term(self, X, Y, B):
        c = self.c
        k_limit = c["blue_modulation.k_limit"][0]
        y_eff = Y + c["blue_modulation.k_offset"][0] + np.abs(X)
        s = self.block_sum(np.where(B > y_eff, np.minimum(B - y_eff, k_limit), 0.0))
        s = np.where(s >= 32.0 * k_limit, 64.0 * k_limit - s, s)
        s = np.minimum(s, c["blue_modulation.k_max_limit"][0] * k_limit)
        return s * c["blue_modulation.k_mul"][0]

    def field(self, xyb, d):
        """Per-block value before scaling and rounding, in FP64."""
        X, Y, B = (xyb[i].astype(np.float64) for i in range(3))
        c = self.c
        mask = self.compute_mask(self.aq_map(self.pre_map(Y), d)) + self.gamma_term(X, Y)
        s0, s1, s2 = c["hf_modulation_strength"]
        strength = min(max(s0 * d, s1), s2)
        hf = (self.hf_sum(Y) * (c["hf_modulation.k_mul_y"][0] * strength)
              + c["hf_modulation.k_offset"][0] + mask)
        out = np.minimum(hf, self.blue_term(X, Y, B) + mask)
        scale = self.k_ac_quant / d
        dampen = min(max(self.schedule("AQ_STRENGTH", d), 0.0), 1.0)
        base = self.schedule("AQ_BASE_LEVEL", d) * scale
        q = np.exp2(out * c["fill_quant_field_scalar.exp_mul"][0])
        return q * (scale * dampen) + (1.0 - dampen) * base


@pytest.fixture(scope="module")
def model(golden):
    return Model(golden)


def planes(X, Y, B, h=64, w=96):
    out = np.empty((3, h, w), np.float32)
    for i, v in enumerate((X, Y, B)):
        out[i] = np.broadcast_to(np.asarray(v, np.float32), (h, w))
    return out


def step_x(a, b, at, h=64, w=96):
    y = np.full((h, w), a, np.float32)
    y[:, at:] = b
    return y


def step_y(a, b, at, h=64, w=96):
    y = np.full((h, w), a, np.float32)
    y[at:, :] = b
    return y


def checker(a, b, h=64, w=96):
    yy, xx = np.mgrid[0:h, 0:w]
    return np.where((yy + xx) % 2 == 0, a, b).astype(np.float32)


def textured(h=64, w=96, seed=3):
    """A checkerboard whose amplitude changes per 4x4 pre-map cell, so the
    nine neighbours of every erosion site are distinct."""
    amp = np.random.default_rng(seed).uniform(0.0, 0.03, (h // 4, w // 4))
    amp = np.kron(amp, np.ones((4, 4)))
    yy, xx = np.mgrid[0:h, 0:w]
    return (0.35 + np.where((yy + xx) % 2 == 0, amp, -amp)).astype(np.float32)


def cases(golden):
    k_limit = np.float32(golden["AQ"]["blue_modulation.k_limit"][0])
    k_off = np.float32(golden["AQ"]["blue_modulation.k_offset"][0])
    y0, x0 = np.float32(0.45), np.float32(0.01)
    y_eff = y0 + k_off + abs(x0)
    return {
        "constant": planes(x0, y0, np.float32(0.40)),
        "constant_dark": planes(np.float32(-0.004), np.float32(0.02), np.float32(0.0)),
        "blue_partial": planes(x0, y0, y_eff + np.float32(0.2) * k_limit),
        "blue_fold": planes(x0, y0, y_eff + np.float32(3.0) * k_limit),
        "step_x_small": planes(x0, step_x(0.45, 0.46, 36), np.float32(0.40)),
        "step_x_large": planes(x0, step_x(0.45, 0.50, 36), np.float32(0.40)),
        "step_x_dark": planes(np.float32(0.0), step_x(0.0, 0.1, 36), np.float32(0.0)),
        "step_x_block_edge": planes(x0, step_x(0.30, 0.60, 40), np.float32(0.20)),
        "step_y_large": planes(x0, step_y(0.45, 0.50, 20), np.float32(0.40)),
        "step_y_dark": planes(np.float32(0.0), step_y(0.1, 0.0, 44), np.float32(0.0)),
        "checker_dark": planes(np.float32(0.0), checker(0.0, 0.1), np.float32(0.0)),
        "checker_small": planes(x0, checker(0.40, 0.42), np.float32(0.30)),
        "textured": planes(x0, textured(), np.float32(0.30)),
    }


# Observed worst relative error over these cases and the photo is 2.0e-7 (the
# float32 chain, dirty_log2f within 2e-6 absolute times k_gam, fast_exp2
# within 4e-7).  The terms the images isolate move the field by 1e-3 to 3e-2
# relative (a step's HF term, the blue boost, the dark-side clamp), so a
# dropped or mistyped term lands far outside 2e-6.
REL_TOL = 2e-6


@pytest.mark.parametrize("d", DISTANCES)
@pytest.mark.parametrize("name", ["constant", "constant_dark", "blue_partial", "blue_fold",
                                  "step_x_small", "step_x_large", "step_x_dark",
                                  "step_x_block_edge", "step_y_large", "step_y_dark",
                                  "checker_dark", "checker_small", "textured"])
def test_float_field_matches_fp64_model(oracle, golden, model, name, d):
    xyb = cases(golden)[name]
    got = oracle.quant_field_float(xyb, d).astype(np.float64)
    want = model.field(xyb, d)
    rel = np.abs(got - want) / want
    assert rel.max() < REL_TOL, (name, d, rel.max(), np.unravel_index(rel.argmax(), rel.shape))


@pytest.mark.parametrize("d", DISTANCES)
def test_constant_image_closed_form(oracle, golden, model, d):
    """Constant planes: diff = 0, so every pixel contributes masking_sqrt(0)
    = 0.25 sqrt(k_log_offset); the pre-map holds 4 of those (16 summed,
    times 0.25), erosion's weights sum to k_total, the block sums 2x2 of
    them, HF is k_offset alone and the gamma term is log2 of one ratio."""
    c = golden["AQ"]
    X, Y, B = 0.01, 0.45, 0.40
    xyb = planes(np.float32(X), np.float32(Y), np.float32(B))
    X, Y = float(np.float32(X)), float(np.float32(Y))
    m0 = 0.25 * math.sqrt(c["masking_sqrt.k_log_offset"][0])
    aq = 4.0 * c["fill_quant_field_scalar.k_total"][0] * 4.0 * m0
    mask = float(model.compute_mask(np.array(aq)))
    iny = Y + c["gamma_modulation.k_bias"][0]
    r = 0.5 * (float(model.ratio(np.array(iny - X), True)) + float(model.ratio(np.array(iny + X), True)))
    mask += c["gamma_modulation.k_gam"][0] * math.log2(r)
    out = mask + c["hf_modulation.k_offset"][0]  # blue term: B below Y, sum 0, so blue = mask
    out = min(out, mask)
    scale = c["K_AC_QUANT"][0] / d
    dampen = model.schedule("AQ_STRENGTH", d)
    want = 2.0 ** (out * c["fill_quant_field_scalar.exp_mul"][0]) * scale * dampen + (
        1.0 - dampen) * model.schedule("AQ_BASE_LEVEL", d) * scale
    got = oracle.quant_field_float(xyb, d)
    assert np.allclose(got, got[0, 0], rtol=0, atol=0), "a constant image gives one value"
    assert abs(float(got[0, 0]) - want) / want < REL_TOL
    assert abs(float(model.field(xyb, d)[0, 0]) - want) / want < 1e-12


@pytest.mark.parametrize("d", DISTANCES)
def test_photo_float_field_matches_fp64_model(oracle, model, d):
    from jixel_amd import XYB_SPEC_FWD
    from jixel_amd.synth import photo_like
    xyb = oracle.xyb_image(photo_like(128, 96, 5), XYB_SPEC_FWD)
    got = oracle.quant_field_float(xyb, d).astype(np.float64)
    want = model.field(xyb, d)
    rel = np.abs(got - want) / want
    assert rel.max() < REL_TOL, rel.max()


@pytest.mark.parametrize("d", DISTANCES)
def test_integer_field_is_the_rounded_float_field(oracle, golden, d):
    """The stored field is clamp(int(q * inv_scale + 0.5), 1, 255)
    (src/adaptive_quant.rs:725-726), on the same float chain."""
    p = oracle.distance_params(d)
    inv_scale = np.float32(1.0 / p["scale"])
    for name, xyb in cases(golden).items():
        fq = oracle.quant_field_float(xyb, d)
        qf = oracle.fill_quant_field(xyb, d, inv_scale)
        want = np.clip(np.floor(fq * inv_scale + np.float32(0.5)), 1, 255).astype(np.uint8)
        assert np.array_equal(qf, want), name


def test_terms_are_isolated(golden, model):
    """The cases exercise what they are named for (a check of the images)."""
    cs = cases(golden)
    k_limit = golden["AQ"]["blue_modulation.k_limit"][0]
    X, Y, B = (cs["blue_partial"][i].astype(np.float64) for i in range(3))
    s = model.blue_term(X, Y, B) / golden["AQ"]["blue_modulation.k_mul"][0]
    assert np.allclose(s, 64 * 0.2 * k_limit, rtol=1e-4)
    X, Y, B = (cs["blue_fold"][i].astype(np.float64) for i in range(3))
    assert np.allclose(model.blue_term(X, Y, B), 0.0, atol=1e-9)
    # the dark side of the step reaches the 0.2 clamp, the bright side does not
    Y = cs["step_x_dark"][1].astype(np.float64)
    P = np.pad(Y, 1, mode="edge")
    base = 0.25 * (P[2:, 1:-1] + P[:-2, 1:-1] + P[1:-1, :-2] + P[1:-1, 2:])
    diff = (model.ratio(Y + model.gamma_offset, False) * (Y - base)) ** 2
    assert diff[0, 35] > 0.2 > diff[0, 36] > 0
    # on the dark checkerboard every dark pixel is clamped, so the clamp
    # reaches the erosion's four smallest
    Y = cs["checker_dark"][1].astype(np.float64)
    P = np.pad(Y, 1, mode="edge")
    base = 0.25 * (P[2:, 1:-1] + P[:-2, 1:-1] + P[1:-1, :-2] + P[1:-1, 2:])
    diff = (model.ratio(Y + model.gamma_offset, False) * (Y - base)) ** 2
    assert (diff[Y == 0] > 0.2).all() and (diff[Y > 0] < 0.2).all()
    # the textured pre-map has nine distinct neighbours at every site
    pre = model.pre_map(cs["textured"][1].astype(np.float64))
    P = np.pad(pre, 1, mode="edge")
    nb = np.stack([P[1 + dy:1 + dy + pre.shape[0], 1 + dx:1 + dx + pre.shape[1]]
                   for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    assert (np.diff(np.sort(nb, axis=0)[:5], axis=0)[:, 1:-1, 1:-1] > 0).all()
    # HF: the large step is clamped at valmin_y, the small one is not
    hf_l = model.hf_sum(cs["step_x_large"][1].astype(np.float64))
    hf_s = model.hf_sum(cs["step_x_small"][1].astype(np.float64))
    assert math.isclose(hf_l[0, 4], 8 * 0.0206, rel_tol=1e-6)
    assert math.isclose(hf_s[0, 4], 8 * float(np.float32(0.46) - np.float32(0.45)), rel_tol=1e-6)
    assert hf_l[0, 3] == 0 and hf_l[0, 5] == 0
    hf_y = model.hf_sum(cs["step_y_large"][1].astype(np.float64))
    assert math.isclose(hf_y[2, 0], 8 * 0.0206, rel_tol=1e-6) and hf_y[1, 0] == 0
    # the step on a block edge has no HF term: pairs never cross a block
    assert model.hf_sum(cs["step_x_block_edge"][1].astype(np.float64)).max() == 0
