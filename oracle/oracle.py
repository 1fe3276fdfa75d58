"""ctypes wrapper of the CPU oracle (oracle/_build/libjxo.so).

TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the benchmark's
CPU-baseline leg use this as the checker.  The product (jixel_amd) never
imports it.  Build with `make -C oracle`.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# JXO_LIB selects another build of the same sources (the benchmark's
# -march=native CPU-baseline build).
_LIB_PATH = os.environ.get("JXO_LIB") or os.path.join(_HERE, "_build", "libjxo.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib, _LIB_PATH
    if _lib is None:
        _LIB_PATH = os.environ.get("JXO_LIB") or _LIB_PATH
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_f32p = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i8p = ctypes.POINTER(ctypes.c_int8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i16p = ctypes.POINTER(ctypes.c_int16)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def _declare(L):
    c = ctypes
    L.jxo_set_fused.argtypes = [c.c_int]
    L.jxo_get_fused.restype = c.c_int
    L.jxo_set_static_tables.argtypes = [_f32p, _f32p]
    L.jxo_fast_tables.argtypes = [c.c_float, c.c_int, _f32p]
    L.jxo_cbrtf.restype = c.c_float
    L.jxo_cbrtf.argtypes = [c.c_float]
    L.jxo_fast_exp2.restype = c.c_float
    L.jxo_fast_exp2.argtypes = [c.c_float]
    L.jxo_dirty_log2f.restype = c.c_float
    L.jxo_dirty_log2f.argtypes = [c.c_float]
    L.jxo_fast_exp.restype = c.c_float
    L.jxo_fast_exp.argtypes = [c.c_float]
    L.jxo_distance_from_quality.restype = c.c_float
    L.jxo_distance_from_quality.argtypes = [c.c_float]
    L.jxo_srgb_lut.argtypes = [_f32p]
    L.jxo_distance_params.argtypes = [c.c_float, _i32p, _f32p, _i32p, _f32p, _u32p, _u32p, _f32p]
    L.jxo_xyb_pixel.argtypes = [_f32p, c.c_int, c.c_float, c.c_float, c.c_float, _f32p]
    L.jxo_xyb_image.argtypes = [_u8p, c.c_size_t, c.c_int, c.c_int, c.c_int, _f32p, c.c_int, _f32p]
    L.jxo_fill_quant_field_image.argtypes = [_f32p, c.c_int, c.c_int, c.c_float, c.c_float, _u8p]
    L.jxo_quant_field_float_image.argtypes = [_f32p, c.c_int, c.c_int, c.c_float, _f32p]
    L.jxo_gaborish_weights.argtypes = [c.c_float, _f32p]
    L.jxo_gaborish_inverse.argtypes = [_f32p, c.c_int, c.c_int, c.c_float]
    L.jxo_forward_transform.restype = c.c_int
    L.jxo_forward_transform.argtypes = [c.c_int, _f32p, c.c_int, c.c_int, c.c_int, c.c_int, _f32p]
    L.jxo_dc_from_coefficients.argtypes = [c.c_int, _f32p, _f32p]
    L.jxo_ac_thresholds.argtypes = [c.c_int, c.c_int, c.c_int, c.c_float, c.c_float, _f32p]
    L.jxo_dequant_level.restype = c.c_float
    L.jxo_dequant_level.argtypes = [c.c_float, c.c_float, c.c_int32]
    L.jxo_forward_quantize.argtypes = [_f32p, c.c_int, c.c_int, _u8p, _u8p, _i8p, _i8p, _f32p,
                                       c.c_void_p, c.POINTER(_i32p), c.POINTER(_f32p),
                                       c.POINTER(_i16p)]
    L.jxo_scan_pos_lut.argtypes = [c.c_int, c.c_int, _u32p]
    L.jxo_natural_scan.argtypes = [c.c_int, c.c_int, _u32p]
    L.jxo_strategy_cost.restype = c.c_float
    L.jxo_strategy_cost.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_float,
                                    c.c_float, _f32p, _f32p, c.c_void_p]
    L.jxo_select_strategy.argtypes = [_f32p, c.c_int, c.c_int, _u8p, _i8p, _i8p, _f32p, c.c_void_p,
                                      _u8p]
    L.jxo_strategy_costs.argtypes = [_f32p, c.c_int, c.c_int, _u8p, _i8p, _i8p, _f32p, c.c_void_p,
                                     _f32p, _f32p, _f32p]
    L.jxo_epf_new.restype = c.c_void_p
    L.jxo_epf_new.argtypes = [_f32p, c.c_int, c.c_int]
    L.jxo_epf_free.argtypes = [c.c_void_p]
    L.jxo_epf_compare.restype = c.c_float
    L.jxo_epf_compare.argtypes = [c.c_void_p, _f32p, c.c_int, _f32p]
    L.jxo_epf_box5.argtypes = [_f32p, c.c_int, c.c_int, _f32p]
    L.jxo_fill_cmap.argtypes = [_f32p, c.c_int, c.c_int, _f32p, c.c_void_p, _i8p, _i8p]
    L.jxo_cmap_tile.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, c.c_int, _f32p, c.c_void_p,
                                c.POINTER(c.c_int), c.POINTER(c.c_int)]
    L.jxo_front_end.argtypes = [_u8p, c.c_size_t, c.c_int, c.c_int, c.c_int, c.c_void_p, _f32p,
                                _i8p, _i8p, c.c_int, _u8p, _u8p, c.POINTER(_i32p),
                                c.POINTER(_f32p), c.POINTER(_i16p), _u8p]
    L.jxo_wc128.restype = _f32p
    L.jxo_dct128.argtypes = [c.c_int, _f32p, _f32p]
    L.jxo_tune_noise.restype = c.c_float
    L.jxo_tune_noise.argtypes = [c.c_uint32, c.c_uint32, c.c_int, c.c_uint32, c.c_int]
    L.jxo_tune_loss_grad.restype = c.c_double
    L.jxo_tune_loss_grad.argtypes = [_f32p, c.c_int, c.c_int, _u8p, c.POINTER(c.c_double), _f32p,
                                     c.c_void_p, c.c_double, c.c_uint32, c.c_uint32,
                                     c.POINTER(c.c_double)]
    L.jxo_chroma_hf_stats.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int,
                                      _f32p, c.POINTER(c.c_int)]
    L.jxo_chroma_hf_strength.restype = c.c_float
    L.jxo_chroma_hf_strength.argtypes = [c.c_float]
    L.jxo_chroma_hf_protection.argtypes = [_f32p, c.c_int, c.c_int, _u8p, c.c_float]
    L.jxo_alpha_fringe_clear.argtypes = [_u8p, c.c_int, c.c_int]
    L.jxo_fill_invisible_plane.argtypes = [_f32p, c.c_size_t, c.c_int, c.c_int, _u8p]
    L.jxo_alpha_block_factors.argtypes = [_u8p, c.c_int, c.c_int, _f32p]
    L.jxo_apply_alpha_factors.argtypes = [_u8p, _u8p, c.c_int, c.c_int, _f32p]
    L.jxo_partition_cost.restype = c.c_float
    L.jxo_partition_cost.argtypes = [_f32p, _f32p, c.c_int]
    L.jxo_choose_merge_partition.restype = c.c_int
    L.jxo_choose_merge_partition.argtypes = [_f32p, _f32p, _f32p, _f32p]
    L.jxo_choose_square32.restype = c.c_int
    L.jxo_choose_square32.argtypes = [_f32p, _f32p, c.c_float, c.c_float]
    L.jxo_aggregate_quant.restype = c.c_int
    L.jxo_aggregate_quant.argtypes = [c.c_int, c.c_uint32, c.c_int, c.c_float]
    L.jxo_aggregate_qac_2x2.restype = c.c_float
    L.jxo_aggregate_qac_2x2.argtypes = [_f32p, c.c_float, c.c_float]
    L.jxo_cmap_factors.argtypes = [c.c_int, c.c_int, c.c_void_p, _f32p]
    L.jxo_plan_small.restype = c.c_float
    L.jxo_plan_small.argtypes = [_f32p, _f32p, _f32p, _f32p, c.c_uint64, c.c_int, c.c_int,
                                 c.c_float, c.POINTER(c.c_int), _u8p, _u8p]
    L.jxo_live_masks.argtypes = [_f32p, _f32p, _f32p, _f32p, c.c_int, c.c_int, _u8p]
    L.jxo_inverse_transform.restype = c.c_int
    L.jxo_inverse_transform.argtypes = [c.c_int, _f32p, _f32p]
    L.jxo_reconstruction_lambda.restype = c.c_float
    L.jxo_reconstruction_lambda.argtypes = [c.c_float]
    L.jxo_ssim_deficit.restype = c.c_float
    L.jxo_ssim_deficit.argtypes = [_f32p, _f32p, c.c_int, c.c_int, c.c_int, c.c_int]
    L.jxo_qf_prices.argtypes = [c.c_int, _u8p, c.c_int, _f32p]
    L.jxo_qf_trellis.argtypes = [c.c_int, _u8p, _i32p, _u8p, _f32p, c.c_int, c.c_float, _f32p, _u8p]
    L.jxo_recon_coeffs.restype = c.c_int
    L.jxo_recon_coeffs.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, _f32p, _f32p]
    L.jxo_recon_cost.restype = c.c_float
    L.jxo_recon_cost.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_float,
                                 _f32p, _f32p, _f32p, c.c_void_p, c.POINTER(c.c_float),
                                 c.POINTER(c.c_float)]
    L.jxo_retile_pass.argtypes = [_f32p, c.c_int, c.c_int, _u8p, _i8p, _i8p, _f32p, c.c_void_p,
                                  _u8p, c.c_int]
    L.jxo_rerank_enable.argtypes = [c.c_int]
    L.jxo_rerank_plan.restype = c.c_float
    L.jxo_rerank_plan.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, _u8p]
    L.jxo_rerank_pass.argtypes = [_f32p, c.c_int, c.c_int, _u8p, _i8p, _i8p, _f32p, c.c_void_p,
                                  _u8p]


def _p(a, t):
    return a.ctypes.data_as(t)


def set_fused(fused):
    lib().jxo_set_fused(1 if fused else 0)


def srgb_lut():
    out = np.zeros(256, np.float32)
    lib().jxo_srgb_lut(_p(out, _f32p))
    return out


def xyb_image(rgb, fwd, order=0):
    """rgb: (H, W, C) uint8 -> (3, H, W) float32."""
    rgb = np.ascontiguousarray(rgb)
    h, w, ch = rgb.shape
    out = np.zeros((3, h, w), np.float32)
    m = np.ascontiguousarray(fwd, np.float32)
    lib().jxo_xyb_image(_p(rgb, _u8p), w * ch, ch, w, h, _p(m, _f32p), order, _p(out, _f32p))
    return out


def fill_quant_field(xyb, distance, inv_scale):
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    qf = np.zeros(((h + 7) // 8, (w + 7) // 8), np.uint8)
    lib().jxo_fill_quant_field_image(_p(xyb, _f32p), w, h, distance, inv_scale, _p(qf, _u8p))
    return qf


def quant_field_float(xyb, distance):
    """Per-block field before scaling by inv_scale and rounding."""
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    fq = np.zeros(((h + 7) // 8, (w + 7) // 8), np.float32)
    lib().jxo_quant_field_float_image(_p(xyb, _f32p), w, h, distance, _p(fq, _f32p))
    return fq


def forward_transform(strategy, plane, px, py):
    plane = np.ascontiguousarray(plane, np.float32)
    h, w = plane.shape
    out = np.zeros(64 * 64, np.float32)
    n = lib().jxo_forward_transform(strategy, _p(plane, _f32p), w, h, px, py, _p(out, _f32p))
    return out[:n]


def natural_scan(cx, cy):
    out = np.zeros(cx * cy * 64, np.uint32)
    lib().jxo_natural_scan(cx, cy, _p(out, _u32p))
    return out


def scan_pos_lut(w, h):
    out = np.zeros(w * h, np.uint32)
    lib().jxo_scan_pos_lut(w, h, _p(out, _u32p))
    return out


def _outputs(xb, yb):
    levels = [np.zeros(yb * xb * 64, np.int32) for _ in range(3)]
    coeffs = [np.zeros(yb * xb * 64, np.float32) for _ in range(3)]
    dc = [np.zeros(yb * xb, np.int16) for _ in range(3)]
    L = (_i32p * 3)(*[_p(a, _i32p) for a in levels])
    C = (_f32p * 3)(*[_p(a, _f32p) for a in coeffs])
    D = (_i16p * 3)(*[_p(a, _i16p) for a in dc])
    return levels, coeffs, dc, L, C, D


def forward_quantize(xyb, cells, qf, ytox, ytob, tables, params):
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    xb, yb = (w + 7) // 8, (h + 7) // 8
    cells = np.ascontiguousarray(cells, np.uint8)
    qf = np.ascontiguousarray(qf, np.uint8)
    ytox = np.ascontiguousarray(ytox, np.int8)
    ytob = np.ascontiguousarray(ytob, np.int8)
    tables = np.ascontiguousarray(tables, np.float32)
    levels, coeffs, dc, L, C, D = _outputs(xb, yb)
    lib().jxo_forward_quantize(_p(xyb, _f32p), w, h, _p(cells, _u8p), _p(qf, _u8p),
                               _p(ytox, _i8p), _p(ytob, _i8p), _p(tables, _f32p),
                               ctypes.byref(params), L, C, D)
    return levels, coeffs, dc


def select_strategy(xyb, qf, ytox, ytob, tables, params):
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    qf = np.array(qf, np.uint8, copy=True)
    cells = np.zeros(qf.shape, np.uint8)
    lib().jxo_select_strategy(_p(xyb, _f32p), w, h, _p(qf, _u8p),
                              _p(np.ascontiguousarray(ytox, np.int8), _i8p),
                              _p(np.ascontiguousarray(ytob, np.int8), _i8p),
                              _p(np.ascontiguousarray(tables, np.float32), _f32p),
                              ctypes.byref(params), _p(cells, _u8p))
    return cells, qf


def plan_small(leaf, wide, tall, square, fixed, w, h, replace_below=float("inf")):
    """plan_small over a w x h grid: (optimum, [(cell, choice), ...] last
    cell first, or None when no plan was saved).  Costs are 64-entry
    arrays indexed y * w + x."""
    arrs = [np.ascontiguousarray(np.asarray(a, np.float32).reshape(64)) for a in
            (leaf, wide, tall, square)]
    n = ctypes.c_int(0)
    cells = np.zeros(64, np.uint8)
    choices = np.zeros(64, np.uint8)
    t = lib().jxo_plan_small(*[_p(a, _f32p) for a in arrs], int(fixed), w, h, replace_below,
                             ctypes.byref(n), _p(cells, _u8p), _p(choices, _u8p))
    return float(t), [(int(cells[i]), int(choices[i])) for i in range(n.value)]


def live_masks(leaf, wide, tall, square, w, h):
    arrs = [np.ascontiguousarray(np.asarray(a, np.float32).reshape(64)) for a in
            (leaf, wide, tall, square)]
    live = np.zeros(64, np.uint8)
    lib().jxo_live_masks(*[_p(a, _f32p) for a in arrs], w, h, _p(live, _u8p))
    return live


def inverse_transform(strategy, coef):
    """The scalar inverse transform of a coefficient block (storage order) to
    its pixel footprint (row-major)."""
    coef = np.ascontiguousarray(coef, np.float32).reshape(-1)
    out = np.zeros(64 * 64, np.float32)
    rc = lib().jxo_inverse_transform(strategy, _p(coef, _f32p), _p(out, _f32p))
    assert rc == 0
    return out[:coef.size]


def reconstruction_lambda(d):
    return float(lib().jxo_reconstruction_lambda(d))


def ssim_deficit(orig, rec, visible=False, w=None, h=None):
    orig = np.ascontiguousarray(orig, np.float32)
    rec = np.ascontiguousarray(rec, np.float32)
    hh, ww = orig.shape
    return float(lib().jxo_ssim_deficit(_p(orig, _f32p), _p(rec, _f32p), ww,
                                        ww if w is None else w, hh if h is None else h,
                                        1 if visible else 0))


def qf_prices(first_left, chain):
    chain = np.ascontiguousarray(chain, np.uint8)
    out = np.zeros((4, 512), np.float32)
    lib().jxo_qf_prices(first_left, _p(chain, _u8p), chain.size, _p(out, _f32p))
    return out


def qf_trellis(first_left, fixed, states, lambda_bits, prices):
    """states: per block None (fixed) or [(q, cost), ...] (up to 4)."""
    n = len(fixed)
    fixed = np.ascontiguousarray(fixed, np.uint8)
    lens = np.zeros(n, np.int32)
    qs = np.zeros((n, 4), np.uint8)
    costs = np.zeros((n, 4), np.float32)
    for i, st in enumerate(states):
        if st:
            lens[i] = len(st)
            for k in range(4):
                q, cst = st[min(k, len(st) - 1)]
                qs[i, k] = q
                costs[i, k] = cst
    chain = np.zeros(n, np.uint8)
    lib().jxo_qf_trellis(first_left, _p(fixed, _u8p), _p(lens, _i32p), _p(qs, _u8p),
                         _p(costs, _f32p), n, lambda_bits,
                         _p(np.ascontiguousarray(prices, np.float32), _f32p), _p(chain, _u8p))
    return chain


def recon_cost(xyb, strategy, px, py, qac, cmap, tables, params):
    """(cost, distortion, rate) of the reconstruction-domain RD model."""
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    cm = np.ascontiguousarray(cmap, np.float32)
    co = np.zeros(3 * 4096, np.float32)
    lib().jxo_recon_coeffs(_p(xyb, _f32p), w, h, strategy, px, py, _p(cm, _f32p), _p(co, _f32p))
    d = ctypes.c_float()
    r = ctypes.c_float()
    cost = lib().jxo_recon_cost(_p(xyb, _f32p), w, h, strategy, px, py, qac, _p(cm, _f32p),
                                _p(co, _f32p), _p(np.ascontiguousarray(tables, np.float32), _f32p),
                                ctypes.byref(params), ctypes.byref(d), ctypes.byref(r))
    return float(cost), float(d.value), float(r.value)


def retile_pass(xyb, qf, ytox, ytob, tables, params, cells, crossing):
    """One coefficient re-tiling pass over `cells` (returned updated)."""
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    cells = np.array(cells, np.uint8, copy=True)
    lib().jxo_retile_pass(_p(xyb, _f32p), w, h, _p(np.ascontiguousarray(qf, np.uint8), _u8p),
                          _p(np.ascontiguousarray(ytox, np.int8), _i8p),
                          _p(np.ascontiguousarray(ytob, np.int8), _i8p),
                          _p(np.ascontiguousarray(tables, np.float32), _f32p),
                          ctypes.byref(params), _p(cells, _u8p), 1 if crossing else 0)
    return cells


def rerank_enable(on):
    """Tests only: the decision with (True) or without the SSIM rerank."""
    lib().jxo_rerank_enable(1 if on else 0)


def rerank_plan(tiles, width, height):
    """best_reconstruction_plan: (cost, 16 tile indices by grid cell, 255 none)
    for tiles [(x, y, w, h, cost), ...]."""
    arr = np.ascontiguousarray(np.asarray(tiles, np.float32).reshape(-1))
    out = np.zeros(16, np.uint8)
    cost = lib().jxo_rerank_plan(_p(arr, _f32p), len(tiles), width, height, _p(out, _u8p))
    return float(cost), out


def rerank_pass(xyb, qf, ytox, ytob, tables, params, cells):
    """The SSIM rerank's downgrade pass over `cells` (returned updated)."""
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    cells = np.array(cells, np.uint8, copy=True)
    lib().jxo_rerank_pass(_p(xyb, _f32p), w, h, _p(np.ascontiguousarray(qf, np.uint8), _u8p),
                          _p(np.ascontiguousarray(ytox, np.int8), _i8p),
                          _p(np.ascontiguousarray(ytob, np.int8), _i8p),
                          _p(np.ascontiguousarray(tables, np.float32), _f32p),
                          ctypes.byref(params), _p(cells, _u8p))
    return cells


def strategy_costs(xyb, qf, ytox, ytob, tables, params):
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    qf = np.ascontiguousarray(qf, np.uint8)
    outs = [np.zeros(qf.shape, np.float32) for _ in range(3)]
    lib().jxo_strategy_costs(_p(xyb, _f32p), w, h, _p(qf, _u8p),
                             _p(np.ascontiguousarray(ytox, np.int8), _i8p),
                             _p(np.ascontiguousarray(ytob, np.int8), _i8p),
                             _p(np.ascontiguousarray(tables, np.float32), _f32p),
                             ctypes.byref(params), *[_p(o, _f32p) for o in outs])
    return outs


def strategy_cost(xyb, strategy, px, py, qac, cmap, tables, params, meta_r=4.0):
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    cm = np.ascontiguousarray(cmap, np.float32)
    return lib().jxo_strategy_cost(_p(xyb, _f32p), w, h, strategy, px, py, qac, meta_r,
                                   _p(cm, _f32p), _p(np.ascontiguousarray(tables, np.float32), _f32p),
                                   ctypes.byref(params))


_GOLDEN = os.path.join(os.path.dirname(_HERE), "tests", "golden", "reference_tables.json")
_static_set = False
_table_cache = {}


def fast_tables(distance, fused=True, arbiter=0):
    """The oracle's own DequantMatrices::new_fast(distance, arbiter) blob
    (src/quant_weights.rs:2294), built with the given fmla flavour.  The
    stored library tables come from the reference-table fixture."""
    global _static_set
    key = (np.float32(distance).item(), bool(fused), int(arbiter))
    if key in _table_cache:
        return _table_cache[key]
    L = lib()
    if not _static_set:
        import json
        with open(_GOLDEN) as f:
            g = json.load(f)
        m8 = np.array(g["DEQUANT_MATRIX_8X8"], np.float32).reshape(-1)
        m16 = np.array(g["DEQUANT_MATRIX_16X8"], np.float32).reshape(-1)
        L.jxo_set_static_tables(_p(m8, _f32p), _p(m16, _f32p))
        _static_set = True
    blob = np.zeros(52224, np.float32)
    old = L.jxo_get_fused()
    L.jxo_set_fused(1 if fused else 0)
    try:
        rc = L.jxo_fast_tables(float(distance), int(arbiter), _p(blob, _f32p))
    finally:
        L.jxo_set_fused(old)
    assert rc == 0
    _table_cache[key] = blob
    return blob


def tables_for(params, tables=None):
    """`tables`, or the oracle's default tables of the frame parameters."""
    if tables is not None:
        return tables
    return fast_tables(params.distance, bool(params.fused), int(params.arbiter))


def front_end(rgb, params, tables=None, threads=1):
    """Whole front end on the CPU.  rgb: (H, W, C) uint8.  Returns the quant
    field, strategy cells, CfL maps, levels, coefficients and DC.  `tables`
    defaults to the oracle's new_fast tables of the frame's distance."""
    tables = tables_for(params, tables)
    rgb = np.ascontiguousarray(rgb)
    h, w, ch = rgb.shape
    xb, yb = (w + 7) // 8, (h + 7) // 8
    qf = np.zeros((yb, xb), np.uint8)
    cells = np.zeros((yb, xb), np.uint8)
    tiles = ((h + 63) // 64, (w + 63) // 64)
    ytox = np.zeros(tiles, np.int8)
    ytob = np.zeros(tiles, np.int8)
    levels, coeffs, dc, L, C, D = _outputs(xb, yb)
    alpha = np.zeros((h, w), np.uint8)
    lib().jxo_front_end(_p(rgb, _u8p), w * ch, ch, w, h, ctypes.byref(params),
                        _p(np.ascontiguousarray(tables, np.float32), _f32p),
                        _p(ytox, _i8p), _p(ytob, _i8p), threads,
                        _p(qf, _u8p), _p(cells, _u8p), L, C, D, _p(alpha, _u8p))
    out = dict(qf=qf, cells=cells, ytox=ytox, ytob=ytob, levels=levels, coeffs=coeffs, dc=dc)
    if ch in (2, 4):
        out["alpha"] = alpha
    return out


def _f32(a, n):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    assert a.size == n
    return a


def partition_cost(children, merges, mask):
    c, m = _f32(children, 4), _f32(merges, 5)
    return float(lib().jxo_partition_cost(_p(c, _f32p), _p(m, _f32p), int(mask)))


def choose_merge_partition(children, baseline, merges, accept):
    c, b, m, a = _f32(children, 4), _f32(baseline, 4), _f32(merges, 5), _f32(accept, 5)
    return int(lib().jxo_choose_merge_partition(_p(c, _f32p), _p(b, _f32p), _p(m, _f32p),
                                                _p(a, _f32p)))


def choose_square32(children, baseline, merge, accept):
    c, b = _f32(children, 4), _f32(baseline, 4)
    return int(lib().jxo_choose_square32(_p(c, _f32p), _p(b, _f32p), float(merge), float(accept)))


def aggregate_quant(max_q, total, covered, distance):
    return int(lib().jxo_aggregate_quant(int(max_q), int(total), int(covered), float(distance)))


def aggregate_qac_2x2(qac, scale, distance):
    q = _f32(qac, 4)
    return float(lib().jxo_aggregate_qac_2x2(_p(q, _f32p), float(scale), float(distance)))


def cmap_factors(ytox, ytob, params):
    out = np.zeros(3, np.float32)
    lib().jxo_cmap_factors(int(ytox), int(ytob), ctypes.byref(params), _p(out, _f32p))
    return out


def alpha_fringe_clear(alpha):
    """alpha_fringe::clear of a (H, W) uint8 plane; returns the cleared copy."""
    a = np.ascontiguousarray(alpha, np.uint8).copy()
    h, w = a.shape
    lib().jxo_alpha_fringe_clear(_p(a, _u8p), w, h)
    return a


def fill_invisible(planes, alpha):
    """fill_invisible of (C, H, W) float32 planes under a (H, W) uint8 alpha."""
    out = np.ascontiguousarray(planes, np.float32).copy()
    a = np.ascontiguousarray(alpha, np.uint8)
    h, w = a.shape
    for c in range(out.shape[0]):
        lib().jxo_fill_invisible_plane(_p(out[c], _f32p), w, w, h, _p(a, _u8p))
    return out


def alpha_block_factors(alpha):
    a = np.ascontiguousarray(alpha, np.uint8)
    h, w = a.shape
    f = np.zeros(((h + 7) // 8, (w + 7) // 8), np.float32)
    lib().jxo_alpha_block_factors(_p(a, _u8p), w, h, _p(f, _f32p))
    return f


def apply_alpha_factors(qf, cells, factors):
    q = np.ascontiguousarray(qf, np.uint8).copy()
    yb, xb = q.shape
    lib().jxo_apply_alpha_factors(_p(q, _u8p), _p(np.ascontiguousarray(cells, np.uint8), _u8p),
                                  xb, yb, _p(np.ascontiguousarray(factors, np.float32), _f32p))
    return q


def fill_cmap(xyb, tables, params):
    """CfL maps of (3, H, W) XYB planes: (ytox, ytob), each ceil(H/64) x ceil(W/64)."""
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    tiles = ((h + 63) // 64, (w + 63) // 64)
    ytox = np.zeros(tiles, np.int8)
    ytob = np.zeros(tiles, np.int8)
    lib().jxo_fill_cmap(_p(xyb, _f32p), w, h, _p(np.ascontiguousarray(tables, np.float32), _f32p),
                        ctypes.byref(params), _p(ytox, _i8p), _p(ytob, _i8p))
    return ytox, ytob


def cmap_tile(xyb, tables, params, tx, ty):
    xyb = np.ascontiguousarray(xyb, np.float32)
    _, h, w = xyb.shape
    x, b = ctypes.c_int(), ctypes.c_int()
    lib().jxo_cmap_tile(_p(xyb, _f32p), w, h, tx, ty,
                        _p(np.ascontiguousarray(tables, np.float32), _f32p), ctypes.byref(params),
                        ctypes.byref(x), ctypes.byref(b))
    return x.value, b.value


def distance_params(distance):
    c = ctypes
    gs, qd = c.c_int32(), c.c_int32()
    scale, scale_dc = c.c_float(), c.c_float()
    xq, it = c.c_uint32(), c.c_uint32()
    steps = np.zeros(3, np.float32)
    lib().jxo_distance_params(distance, c.byref(gs), c.byref(scale), c.byref(qd),
                              c.byref(scale_dc), c.byref(xq), c.byref(it), _p(steps, _f32p))
    return dict(global_scale=gs.value, scale=scale.value, quant_dc=qd.value,
                scale_dc=scale_dc.value, x_qm_scale=xq.value, epf_iters=it.value,
                dc_step=steps)


def scalar(name, x):
    return getattr(lib(), "jxo_" + name)(x)


def gaborish_weights(mul):
    w = np.zeros(6, np.float32)
    lib().jxo_gaborish_weights(mul, _p(w, _f32p))
    return w


def gaborish_inverse(xyb, mul=0.9908511):
    """xyb: (3, H, W) float32; returns the filtered copy."""
    out = np.ascontiguousarray(xyb, np.float32).copy()
    _, h, w = out.shape
    lib().jxo_gaborish_inverse(_p(out, _f32p), w, h, mul)
    return out


def epf_compare(y_ref, y_cand, wide=False):
    """(score, terms[6]) of LumaProxy::new(y_ref).compare(y_cand)."""
    y_ref = np.ascontiguousarray(y_ref, np.float32)
    y_cand = np.ascontiguousarray(y_cand, np.float32)
    h, w = y_ref.shape
    L = lib()
    p = L.jxo_epf_new(_p(y_ref, _f32p), w, h)
    try:
        terms = np.zeros(6, np.float32)
        score = L.jxo_epf_compare(p, _p(y_cand, _f32p), 1 if wide else 0, _p(terms, _f32p))
        return float(np.float32(score)), terms
    finally:
        L.jxo_epf_free(p)


def epf_box5(src):
    src = np.ascontiguousarray(src, np.float32)
    h, w = src.shape
    out = np.zeros_like(src)
    lib().jxo_epf_box5(_p(src, _f32p), w, h, _p(out, _f32p))
    return out


def chroma_hf_stats(xyb, px, py, w, h):
    """(chroma_sum, chroma_grad, luma_grad, edges) of a block of (3, H, W) XYB."""
    x = np.ascontiguousarray(xyb, np.float32)
    out = np.zeros(3, np.float32)
    e = ctypes.c_int()
    lib().jxo_chroma_hf_stats(_p(x, _f32p), x.shape[2], x.shape[1], px, py, w, h, _p(out, _f32p),
                              ctypes.byref(e))
    return float(out[0]), float(out[1]), float(out[2]), e.value


def chroma_hf_strength(distance):
    return float(lib().jxo_chroma_hf_strength(float(distance)))


def chroma_hf_protection(xyb, qf, distance):
    x = np.ascontiguousarray(xyb, np.float32)
    q = np.ascontiguousarray(qf, np.uint8).copy()
    lib().jxo_chroma_hf_protection(_p(x, _f32p), x.shape[2], x.shape[1], _p(q, _u8p),
                                   float(distance))
    return q


def tune_loss_grad(xyb, cells, theta, tables, params, lam, seed, step, with_grad=True):
    """FP64 tuning loss and per-block gradient of (3, H, W) XYB."""
    x = np.ascontiguousarray(xyb, np.float32)
    th = np.ascontiguousarray(theta, np.float64)
    g = np.zeros_like(th) if with_grad else None
    dp = ctypes.POINTER(ctypes.c_double)
    loss = lib().jxo_tune_loss_grad(_p(x, _f32p), x.shape[2], x.shape[1],
                                    _p(np.ascontiguousarray(cells, np.uint8), _u8p),
                                    th.ctypes.data_as(dp),
                                    _p(np.ascontiguousarray(tables, np.float32), _f32p),
                                    ctypes.byref(params), float(lam), int(seed), int(step),
                                    g.ctypes.data_as(dp) if g is not None else None)
    return loss, g


def tune_noise(seed, step, c, first, k):
    return float(lib().jxo_tune_noise(seed, step, c, first, k))


def wc128():
    p = lib().jxo_wc128()
    return np.array([p[i] for i in range(64)], np.float32)


def dct128(kind, block):
    """128-family transform of a footprint ((128, 128), (128, 64) or (64, 128)
    rows x cols for kind 0, 1, 2) in the reference's storage order."""
    b = np.ascontiguousarray(block, np.float32)
    out = np.zeros(b.size, np.float32)
    lib().jxo_dct128(int(kind), _p(b, _f32p), _p(out, _f32p))
    return out
