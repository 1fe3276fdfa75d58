"""ctypes binding of libjxg.so, the C ABI declared in include/jxg.h."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjxg.so")

JXG_OK = 0
ERRORS = {
    -1: "invalid argument",
    -2: "empty image",
    -3: "dimension too large",
    -4: "invalid distance",
    -5: "allocation failed",
    -6: "device error",
    -7: "no gfx950 device",
    -8: "unsupported speed tier",
}

EXPORTED = [
    "jxg_qt_size", "jxg_qt_offset", "jxg_qt_kind_of", "jxg_distance_from_quality",
    "jxg_frame_params_init", "jxg_srgb_lut", "jxg_default_quant_tables", "jxg_plane_geometry",
    "jxg_spec_xyb_matrix",
    "jxg_ctx_create", "jxg_ctx_destroy", "jxg_ctx_synchronize", "jxg_last_error",
    "jxg_ctx_set_tables", "jxg_xyb_dev", "jxg_fill_quant_field_dev", "jxg_strategy_costs_dev",
    "jxg_select_strategy_dev", "jxg_forward_quantize_dev", "jxg_front_end_dev", "jxg_front_end",
    "jxg_quantize_status", "jxg_gaborish_inverse_dev", "jxg_fill_cmap_dev",
    "jxg_dct64_image_dev", "jxg_epf_proxy_new", "jxg_epf_proxy_destroy", "jxg_epf_proxy_compare",
    "jxg_alpha_fringe_clear_dev", "jxg_fill_invisible_dev", "jxg_alpha_block_factors_dev",
    "jxg_apply_alpha_factors_dev", "jxg_alpha_plane", "jxg_chroma_hf_protection_dev",
    "jxg_strategy_cost_sites_dev", "jxg_batch_create", "jxg_batch_destroy", "jxg_batch_run",
    "jxg_batch_launch", "jxg_batch_status", "jxg_batch_autotune", "jxg_batch_config",
    "jxg_qf_tuner_create", "jxg_qf_tuner_destroy", "jxg_qf_tuner_excluded",
    "jxg_qf_tuner_loss_grad", "jxg_qf_tuner_run", "jxg_dct128_image_dev",
    "jxg_fill_quant_field_noalloc_dev", "jxg_ctx_profile", "jxg_ctx_profile_read",
]


class FrameParams(ctypes.Structure):
    """Mirror of jxg_frame_params (include/jxg.h)."""
    _fields_ = [
        ("distance", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("scale_dc", ctypes.c_float),
        ("dc_step", ctypes.c_float * 3),
        ("ytob_dc", ctypes.c_int32),
        ("cfl_base_x", ctypes.c_float),
        ("cfl_base_b", ctypes.c_float),
        ("cfl_color_factor", ctypes.c_float),
        ("x_qm_scale", ctypes.c_uint32),
        ("b_qm_scale", ctypes.c_uint32),
        ("dequant_one", ctypes.c_float * 3),
        ("dequant_numerator", ctypes.c_float),
        ("fused", ctypes.c_int32),
        ("speed", ctypes.c_int32),
        ("global_scale", ctypes.c_int32),
        ("quant_dc", ctypes.c_int32),
        ("epf_iters", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("channel_weight", ctypes.c_float * 3),
        ("arbiter", ctypes.c_int32),
        ("xyb_order", ctypes.c_int32),
        ("xyb_fwd", ctypes.c_float * 9),
    ]


class BatchImage(ctypes.Structure):
    """Mirror of jxg_batch_image (include/jxg.h)."""
    _fields_ = [
        ("d_rgb", ctypes.c_void_p),
        ("row_bytes", ctypes.c_size_t),
        ("channels", ctypes.c_int32),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("params", FrameParams),
        ("d_ytox", ctypes.c_void_p),
        ("d_ytob", ctypes.c_void_p),
        ("d_qf", ctypes.c_void_p),
        ("d_cells", ctypes.c_void_p),
        ("d_levels", ctypes.c_void_p * 3),
        ("d_coeffs", ctypes.c_void_p * 3),
        ("d_dc", ctypes.c_void_p * 3),
    ]


class JxgError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"jxg error {code} ({ERRORS.get(code, '?')}): {message}")
        self.code = code


_lib = None


def lib():
    """Loads libjxg.so; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m jixel_amd.build` "
            "(the front end has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    _declare(L)
    _lib = L
    return L


def _declare(L):
    c = ctypes
    vp = c.c_void_p
    L.jxg_qt_size.restype = c.c_uint32
    L.jxg_qt_size.argtypes = [c.c_int]
    L.jxg_qt_offset.restype = c.c_uint32
    L.jxg_qt_offset.argtypes = [c.c_int]
    L.jxg_qt_kind_of.argtypes = [c.c_int]
    L.jxg_distance_from_quality.restype = c.c_float
    L.jxg_distance_from_quality.argtypes = [c.c_float]
    L.jxg_frame_params_init.argtypes = [c.c_float, c.c_int, c.c_int, c.POINTER(FrameParams)]
    L.jxg_srgb_lut.argtypes = [vp]
    L.jxg_default_quant_tables.argtypes = [c.c_float, c.c_int, c.c_int, vp]
    L.jxg_spec_xyb_matrix.argtypes = [vp]
    L.jxg_spec_xyb_matrix.restype = None
    L.jxg_plane_geometry.argtypes = [c.c_uint32, c.c_uint32, c.POINTER(c.c_uint32),
                                     c.POINTER(c.c_uint32)]
    L.jxg_plane_geometry.restype = None
    L.jxg_ctx_create.argtypes = [c.c_int, vp, c.POINTER(vp)]
    L.jxg_ctx_destroy.argtypes = [vp]
    L.jxg_ctx_destroy.restype = None
    L.jxg_ctx_synchronize.argtypes = [vp]
    L.jxg_last_error.restype = c.c_char_p
    L.jxg_ctx_set_tables.argtypes = [vp, vp]
    L.jxg_xyb_dev.argtypes = [vp, vp, c.c_size_t, c.c_int, c.c_uint32, c.c_uint32, vp, vp, c.c_int,
                              c.c_int, vp, c.c_uint32, c.c_uint32]
    L.jxg_fill_quant_field_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                           c.c_float, c.c_float, c.c_int, vp]
    L.jxg_fill_quant_field_noalloc_dev.argtypes = L.jxg_fill_quant_field_dev.argtypes
    L.jxg_ctx_profile.argtypes = [vp, c.c_int]
    L.jxg_ctx_profile_read.argtypes = [vp, vp, vp]
    L.jxg_strategy_costs_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                         vp, vp, vp, c.POINTER(FrameParams), vp, vp, vp]
    L.jxg_select_strategy_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                          vp, vp, vp, c.POINTER(FrameParams), vp]
    L.jxg_forward_quantize_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                           vp, vp, vp, vp, c.POINTER(FrameParams), vp, vp, vp]
    L.jxg_front_end_dev.argtypes = [vp, vp, c.c_size_t, c.c_int, c.c_uint32, c.c_uint32,
                                    c.POINTER(FrameParams), vp, vp, vp, vp, vp, vp, vp]
    L.jxg_front_end.argtypes = [vp, vp, c.c_size_t, c.c_int, c.c_uint32, c.c_uint32,
                                c.POINTER(FrameParams), vp, vp, vp, vp, vp, vp, vp]
    L.jxg_quantize_status.argtypes = [vp, vp]
    L.jxg_dct64_image_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                      c.c_int, c.c_int, vp]
    L.jxg_epf_proxy_new.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.POINTER(vp)]
    L.jxg_epf_proxy_destroy.argtypes = [vp]
    L.jxg_epf_proxy_destroy.restype = None
    L.jxg_epf_proxy_compare.argtypes = [vp, vp, c.c_uint32, c.POINTER(c.c_float), vp]
    L.jxg_fill_cmap_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                    c.POINTER(FrameParams), vp, vp]
    L.jxg_gaborish_inverse_dev.argtypes = [vp, vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                           c.c_float, c.c_int]
    L.jxg_alpha_fringe_clear_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32]
    L.jxg_fill_invisible_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32, vp]
    L.jxg_alpha_block_factors_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, vp]
    L.jxg_apply_alpha_factors_dev.argtypes = [vp, vp, vp, c.c_uint32, c.c_uint32, vp]
    L.jxg_alpha_plane.argtypes = [vp, vp, c.c_size_t, c.c_int]
    L.jxg_dct128_image_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                       c.c_int, c.c_int, vp]
    L.jxg_qf_tuner_create.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32, vp,
                                      c.POINTER(FrameParams), c.c_float, c.c_uint32, c.c_uint32,
                                      c.POINTER(vp)]
    L.jxg_qf_tuner_destroy.argtypes = [vp]
    L.jxg_qf_tuner_destroy.restype = None
    L.jxg_qf_tuner_excluded.argtypes = [vp, c.POINTER(c.c_uint32)]
    L.jxg_qf_tuner_loss_grad.argtypes = [vp, vp, c.c_uint32, vp, c.POINTER(c.c_float)]
    L.jxg_qf_tuner_run.argtypes = [vp, vp, c.c_uint32, c.c_float, vp, vp, vp]
    L.jxg_batch_create.argtypes = [vp, c.c_uint32, c.POINTER(BatchImage), c.POINTER(vp)]
    L.jxg_batch_destroy.argtypes = [vp]
    L.jxg_batch_destroy.restype = None
    L.jxg_batch_run.argtypes = [vp]
    L.jxg_batch_launch.argtypes = [vp]
    L.jxg_batch_status.argtypes = [vp, vp]
    L.jxg_batch_autotune.argtypes = [vp, c.c_int, c.c_char_p, c.c_size_t]
    L.jxg_batch_config.argtypes = [vp, vp]
    L.jxg_strategy_cost_sites_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32,
                                              c.c_uint32, vp, vp, c.POINTER(FrameParams), c.c_int,
                                              c.c_uint32, vp, vp, c.c_float, vp]
    L.jxg_chroma_hf_protection_dev.argtypes = [vp, vp, c.c_uint32, c.c_uint32, c.c_uint32,
                                               c.c_uint32, c.c_float, c.c_int, vp]


def check(rc):
    if rc != JXG_OK:
        msg = lib().jxg_last_error()
        raise JxgError(rc, msg.decode() if msg else "")
    return rc
