"""MI355X (gfx950) VarDCT encoder front end for jixel.

The product is libjxg.so (C ABI, include/jxg.h): hand-written HIP kernels for
sRGB -> XYB, the adaptive quant field, the AC-strategy search and forward
DCT + quantization.  This package is the Python host mirror of the
reference's configuration surface (EncodeConfig / Speed /
distance_from_quality, src/encode_image.rs:326-1337) and a thin driver over
the C ABI.  torch is used only to own device memory and streams.

There is no CPU fallback: every entry point raises when libjxg.so or a
gfx950 device is missing.
"""
import ctypes
import enum
from dataclasses import dataclass, field, replace

import numpy as np

from . import _lib
from ._lib import FrameParams, JxgError, check, lib

__all__ = [
    "Speed", "EncodeConfig", "FrameParams", "JxgError", "FrontEnd", "FrontEndOutput", "LumaProxy",
    "Batch", "QfTuner",
    "distance_from_quality", "frame_params", "default_quant_tables", "srgb_lut",
    "plane_geometry", "qt_offset", "qt_size", "encode_front_end", "XYB_SPEC_FWD",
    "STRATEGY",
]


class Speed(enum.IntEnum):
    """Encoder speed tiers (src/encode_image.rs Speed, src/effort.rs:427)."""
    ULTRAFAST = 0
    FASTEST = 1
    FAST = 2
    MEDIUM = 3
    MEDIUMSLOW = 4
    SLOW = 5
    EXTRASLOW = 6


# Raw strategy codes (src/dc_group_data.rs:32-56).
STRATEGY = dict(DCT=0, DCT16X8=1, DCT8X16=2, DCT16X16=3, DCT32X32=4, DCT4X4=5, DCT4X8=6,
                DCT8X4=7, DCT32X16=8, DCT16X32=9, AFV0=10, AFV1=11, AFV2=12, AFV3=13,
                DCT64X64=14, DCT64X32=15, DCT32X64=16, IDENTITY=17, DCT2X2=18, DCT128X128=19,
                DCT128X64=20, DCT64X128=21, DCT32X8=22, DCT8X32=23)

# gaborize's multiplier (src/frame/quantization.rs:243).
GABORISH_MUL = 0.9908511


def distance_from_quality(quality):
    """distance_from_quality (src/encode_image.rs:1321-1337); NaN is rejected."""
    if quality != quality:
        raise ValueError("quality must not be NaN")
    return float(lib().jxg_distance_from_quality(float(quality)))


@dataclass(frozen=True)
class EncodeConfig:
    """The front end's view of jixel's EncodeConfig (src/encode_image.rs:326-359).

    `distance` defaults to 1.0 and `speed` to Fast, as the reference's
    Default does (:682-740).  `fused` selects the reference build's fmla
    semantics (aarch64 / FMA-enabled builds fuse, generic x86-64 does not);
    `xyb_order` selects the scalar XYB kernel or the x86-64 AVX2 one.
    `fill_invisible` (on by default, :693) fills the colour under alpha 0
    and scales the quant field by the alpha block factors; `lossy_alpha`
    (alpha_distance unset or > 0) clears the alpha fringe first.
    """
    distance: float = 1.0
    speed: Speed = Speed.FAST
    fused: bool = True
    xyb_order: int = 0
    fill_invisible: bool = True
    lossy_alpha: bool = True
    arbiter: int = 0  # ARBITER_SSIMULACRA2 (src/encode_image.rs:717)

    def with_fill_invisible(self, on):
        return replace(self, fill_invisible=bool(on))

    def with_quality(self, quality):
        return replace(self, distance=distance_from_quality(quality))

    def with_distance(self, distance):
        if not (distance > 0.0) or distance != distance:
            raise JxgError(-4, f"distance {distance}")
        return replace(self, distance=float(distance))

    def with_speed(self, speed):
        return replace(self, speed=Speed(speed))

    def with_arbiter(self, arbiter):
        if int(arbiter) not in (ARBITER_SSIMULACRA2, ARBITER_CVVDP):
            raise JxgError(-1, f"arbiter {arbiter}")
        return replace(self, arbiter=int(arbiter))


def frame_params(config: EncodeConfig) -> FrameParams:
    """compute_distance_params + EncodingContext defaults for `config`."""
    p = FrameParams()
    check(lib().jxg_frame_params_init(float(config.distance), int(config.speed),
                                      1 if config.fused else 0, ctypes.byref(p)))
    p.flags = ((FLAG_FILL_INVISIBLE if config.fill_invisible else 0) |
               (FLAG_LOSSY_ALPHA if config.lossy_alpha else 0))
    p.arbiter = int(config.arbiter)
    p.xyb_order = int(config.xyb_order)
    return p


def qt_size(kind):
    return int(lib().jxg_qt_size(kind))


def qt_offset(kind):
    return int(lib().jxg_qt_offset(kind))


QT_BLOB_FLOATS = 52224


ARBITER_SSIMULACRA2 = 0  # JXG_ARBITER_SSIMULACRA2 (the reference's default)
ARBITER_CVVDP = 1        # JXG_ARBITER_CVVDP


def default_quant_tables(distance, fused=True, arbiter=ARBITER_SSIMULACRA2):
    """DequantMatrices::new_fast(distance, arbiter) (src/quant_weights.rs:2294):
    the tables the device stages use by default for a frame at `distance`."""
    blob = np.zeros(QT_BLOB_FLOATS, np.float32)
    check(lib().jxg_default_quant_tables(float(distance), int(arbiter), 1 if fused else 0,
                                         blob.ctypes.data))
    return blob


def srgb_lut():
    out = np.zeros(256, np.float32)
    lib().jxg_srgb_lut(out.ctypes.data)
    return out


def plane_geometry(width, height):
    s, h = ctypes.c_uint32(), ctypes.c_uint32()
    lib().jxg_plane_geometry(width, height, ctypes.byref(s), ctypes.byref(h))
    return s.value, h.value


def _spec_fwd():
    # XybMatrix::SPEC (src/xyb.rs:35-47,81-96) in f32 arithmetic.
    f = np.float32
    m00, m02 = f(0.30), f(0.078)
    m01 = f(f(1.0) - m02) - m00
    m10, m12 = f(0.23), f(0.078)
    m11 = f(f(1.0) - m12) - m10
    b_bias = f(0.55180986)
    ratio = f(0.24342269) / f(f(0.24342269) + f(0.20476745))
    m20 = f(ratio * f(f(1.0) - b_bias))
    m21 = f(f(f(1.0) - ratio) * f(f(1.0) - b_bias))
    m22 = f(f(f(1.0) - m20) - m21)
    return np.array([m00, m01, m02, m10, m11, m12, m20, m21, m22], np.float32)


XYB_SPEC_FWD = _spec_fwd()


FLAG_FILL_INVISIBLE = 1  # JXG_FLAG_FILL_INVISIBLE
FLAG_LOSSY_ALPHA = 2     # JXG_FLAG_LOSSY_ALPHA
FLAG_CHROMA_HF = 4       # JXG_FLAG_CHROMA_HF


@dataclass
class FrontEndOutput:
    """Per-image front-end results in the layouts of include/jxg.h."""
    qf: np.ndarray      # (yblocks, xblocks) uint8, final quant field
    cells: np.ndarray   # (yblocks, xblocks) uint8 strategy cells
    ytox: np.ndarray    # (ytiles, xtiles) int8 CfL map, Y to X
    ytob: np.ndarray    # (ytiles, xtiles) int8 CfL map, Y to B
    levels: list        # 3 x int32 [yblocks*xblocks*64]
    coeffs: list        # 3 x float32 [yblocks*xblocks*64]
    dc: list            # 3 x int16 [yblocks*xblocks]
    alpha: np.ndarray = None  # (H, W) uint8 alpha plane as the stages used it


def _dev_ptr(t, device, dtype, numel, what):
    """The device address of torch tensor `t` once it is checked to be what
    the C-ABI entry point will read or write: a tensor on `device` of `dtype`,
    contiguous, with at least `numel` elements.  JxgError (invalid argument)
    otherwise, before any launch: the library itself checks only for null
    pointers and plane geometry."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise JxgError(-1, f"{what}: expected a torch tensor, got {type(t).__name__}")
    if not t.is_cuda or t.device != device:
        raise JxgError(-1, f"{what}: expected a tensor on {device}, got {t.device}")
    if t.dtype != dtype:
        raise JxgError(-1, f"{what}: expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise JxgError(-1, f"{what}: expected a contiguous tensor")
    if t.numel() < numel:
        raise JxgError(-1, f"{what}: {t.numel()} elements, the stage needs {numel}")
    return ctypes.c_void_p(t.data_ptr())


class FrontEnd:
    """A device context (jxg_ctx) plus torch-owned buffers for its stages."""

    def __init__(self, device=0, stream=None, tables=None):
        import torch
        self.torch = torch
        if not torch.cuda.is_available():
            raise JxgError(-7, "no GPU visible to torch")
        self.device = torch.device("cuda", device)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        h = ctypes.c_void_p()
        check(lib().jxg_ctx_create(device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)))
        self.ctx = h
        if tables is not None:
            self.set_tables(tables)

    def close(self):
        if getattr(self, "ctx", None):
            lib().jxg_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tables(self, blob):
        blob = np.ascontiguousarray(blob, np.float32)
        assert blob.size == QT_BLOB_FLOATS
        check(lib().jxg_ctx_set_tables(self.ctx, blob.ctypes.data))

    def synchronize(self):
        check(lib().jxg_ctx_synchronize(self.ctx))

    PROFILE_STAGES = ("xyb", "alpha", "cmap", "aq_pre", "aq_block", "select", "alpha_qf",
                      "quantize", "quantize_xt", "fill", "retile", "refine", "rerank")  # JXG_PROF_* order ("fill" is inside "alpha")

    def profile(self, enable=True):
        """Turns the library's per-kernel event timing on (and resets it) or off."""
        check(lib().jxg_ctx_profile(self.ctx, 1 if enable else 0))

    def profile_read(self):
        """{stage: (mean ms per timed launch, launches)} of the timed front-end calls."""
        tot = np.zeros(len(self.PROFILE_STAGES), np.float64)
        cnt = np.zeros(len(self.PROFILE_STAGES), np.uint32)
        check(lib().jxg_ctx_profile_read(self.ctx, tot.ctypes.data, cnt.ctypes.data))
        return {k: (float(t / n) if n else 0.0, int(n))
                for k, t, n in zip(self.PROFILE_STAGES, tot, cnt)}

    # ---- checked device pointers ----
    def _buf(self, t, dtype, numel, what):
        return _dev_ptr(t, self.device, dtype, numel, what)

    def _planes(self, xyb, width, height, what="xyb"):
        """(3, padded_h, stride) float32 planes: pointer, stride, padded_h."""
        t = self.torch
        if not isinstance(xyb, t.Tensor) or xyb.dim() != 3 or xyb.shape[0] != 3:
            raise JxgError(-1, f"{what}: expected (3, padded_h, stride) planes")
        return (self._buf(xyb, t.float32, 0, what), int(xyb.shape[2]), int(xyb.shape[1]))

    def _image(self, rgb, width, height, what="rgb"):
        """(H, W, C) uint8 image with whole-pixel rows (a row stride of its
        own is allowed): pointer, row bytes, channels."""
        t = self.torch
        if not isinstance(rgb, t.Tensor) or rgb.dim() != 3:
            raise JxgError(-1, f"{what}: expected an (H, W, C) tensor")
        if not rgb.is_cuda or rgb.device != self.device:
            raise JxgError(-1, f"{what}: expected a tensor on {self.device}, got {rgb.device}")
        if rgb.dtype != t.uint8:
            raise JxgError(-1, f"{what}: expected torch.uint8, got {rgb.dtype}")
        h, w, ch = (int(v) for v in rgb.shape)
        if h < height or w < width or ch not in (1, 2, 3, 4):
            raise JxgError(-1, f"{what}: shape {tuple(rgb.shape)} for a {width}x{height} image")
        if rgb.stride(2) != 1 or rgb.stride(1) != ch or rgb.stride(0) < w * ch:
            raise JxgError(-1, f"{what}: pixels must be contiguous (strides {rgb.stride()})")
        return ctypes.c_void_p(rgb.data_ptr()), rgb.stride(0), ch

    def _blocks(self, width, height):
        return ((width + 7) // 8) * ((height + 7) // 8)

    def _tiles(self, width, height):
        return ((width + 63) // 64) * ((height + 63) // 64)

    def _out3(self, ts, dtype, numel, what):
        if len(ts) != 3:
            raise JxgError(-1, f"{what}: expected 3 tensors")
        return (ctypes.c_void_p * 3)(*[self._buf(x, dtype, numel, f"{what}[{c}]").value
                                       for c, x in enumerate(ts)])

    def _outputs_arrays(self, o, width, height):
        t, nb = self.torch, self._blocks(width, height)
        return (self._out3(o["levels"], t.int32, nb * 64, "levels"),
                self._out3(o["coeffs"], t.float32, nb * 64, "coeffs"),
                self._out3(o["dc"], t.int16, nb, "dc"))

    # ---- buffers ----
    def empty(self, shape, dtype):
        return self.torch.empty(shape, dtype=dtype, device=self.device)

    def xyb_buffer(self, width, height):
        stride, ph = plane_geometry(width, height)
        return self.empty((3, ph, stride), self.torch.float32)

    def outputs(self, width, height):
        t = self.torch
        xb, yb = (width + 7) // 8, (height + 7) // 8
        return dict(
            levels=[self.empty(yb * xb * 64, t.int32) for _ in range(3)],
            coeffs=[self.empty(yb * xb * 64, t.float32) for _ in range(3)],
            dc=[self.empty(yb * xb, t.int16) for _ in range(3)],
        )

    # ---- stages ----
    def xyb(self, rgb, width, height, fwd=None, lut=None, order=0, fused=True, out=None):
        """rgb: (H, W, C) uint8 device tensor -> (3, padded_h, stride) XYB."""
        fwd = XYB_SPEC_FWD if fwd is None else np.ascontiguousarray(fwd, np.float32)
        lut = srgb_lut() if lut is None else np.ascontiguousarray(lut, np.float32)
        out = self.xyb_buffer(width, height) if out is None else out
        prgb, row_bytes, ch = self._image(rgb, width, height)
        pout, s, ph = self._planes(out, width, height, "out")
        check(lib().jxg_xyb_dev(self.ctx, prgb, row_bytes, ch, width, height, fwd.ctypes.data,
                                lut.ctypes.data, order, 1 if fused else 0, pout, s, ph))
        return out

    def fill_quant_field(self, xyb, width, height, distance, inv_scale, fused=True, out=None):
        xb, yb = (width + 7) // 8, (height + 7) // 8
        out = self.empty((yb, xb), self.torch.uint8) if out is None else out
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_fill_quant_field_dev(self.ctx, px, s, ph, width, height, float(distance),
                                             float(inv_scale), 1 if fused else 0,
                                             self._buf(out, self.torch.uint8,
                                                       self._blocks(width, height), "out")))
        return out

    def _cmaps(self, width, height, ytox, ytob):
        t = self.torch
        n = ((height + 63) // 64, (width + 63) // 64)
        if ytox is None:
            ytox = t.zeros(n, dtype=t.int8, device=self.device)
        if ytob is None:
            ytob = t.zeros(n, dtype=t.int8, device=self.device)
        return ytox, ytob

    def strategy_costs(self, xyb, width, height, qf, params, ytox=None, ytob=None):
        t = self.torch
        ytox, ytob = self._cmaps(width, height, ytox, ytob)
        nb, nt = self._blocks(width, height), self._tiles(width, height)
        outs = [self.empty(qf.shape, t.float32) for _ in range(3)]
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_strategy_costs_dev(self.ctx, px, s, ph, width, height,
                                           self._buf(qf, t.uint8, nb, "qf"),
                                           self._buf(ytox, t.int8, nt, "ytox"),
                                           self._buf(ytob, t.int8, nt, "ytob"),
                                           ctypes.byref(params),
                                           *[self._buf(o, t.float32, nb, "costs") for o in outs]))
        return outs

    def select_strategy(self, xyb, width, height, qf, params, ytox=None, ytob=None, cells=None):
        """Writes strategy cells; updates `qf` in place (adjust_quant_field)."""
        ytox, ytob = self._cmaps(width, height, ytox, ytob)
        t, nb, nt = self.torch, self._blocks(width, height), self._tiles(width, height)
        cells = self.empty(qf.shape, t.uint8) if cells is None else cells
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_select_strategy_dev(self.ctx, px, s, ph, width, height,
                                            self._buf(qf, t.uint8, nb, "qf"),
                                            self._buf(ytox, t.int8, nt, "ytox"),
                                            self._buf(ytob, t.int8, nt, "ytob"),
                                            ctypes.byref(params), self._buf(cells, t.uint8, nb, "cells")))
        return cells

    def forward_quantize(self, xyb, width, height, cells, qf, params, ytox=None, ytob=None,
                         outputs=None):
        ytox, ytob = self._cmaps(width, height, ytox, ytob)
        o = self.outputs(width, height) if outputs is None else outputs
        t, nb, nt = self.torch, self._blocks(width, height), self._tiles(width, height)
        px, s, ph = self._planes(xyb, width, height)
        lv, co, dc = self._outputs_arrays(o, width, height)
        check(lib().jxg_forward_quantize_dev(self.ctx, px, s, ph, width, height,
                                             self._buf(cells, t.uint8, nb, "cells"),
                                             self._buf(qf, t.uint8, nb, "qf"),
                                             self._buf(ytox, t.int8, nt, "ytox"),
                                             self._buf(ytob, t.int8, nt, "ytob"),
                                             ctypes.byref(params), lv, co, dc))
        return o

    def gaborish_inverse(self, xyb, width, height, mul=GABORISH_MUL, fused=True, out=None):
        """Gaborish pre-inversion of (3, padded_h, stride) planes; in place
        unless `out` (same geometry) is given."""
        out = xyb if out is None else out
        px, s, ph = self._planes(xyb, width, height)
        if tuple(out.shape) != tuple(xyb.shape):
            raise JxgError(-1, f"out: shape {tuple(out.shape)} != {tuple(xyb.shape)}")
        po, _, _ = self._planes(out, width, height, "out")
        check(lib().jxg_gaborish_inverse_dev(self.ctx, px, po, s, ph, width, height, float(mul),
                                             1 if fused else 0))
        return out

    def dct64_image(self, xyb, width, height, method=0, fused=True, out=None):
        """DCT64X64 of every 64x64 block (method 0: FP32 butterflies in the
        reference's order, 1: split-bf16 x3 on the matrix cores, 2: the
        three-term split with six products).  Returns 3
        float tensors of 64 * (8 ceil(W/64)) * (8 ceil(H/64)) coefficients in
        the quantize-stage chunk layout with xblocks = 8 ceil(W/64)."""
        n = 64 * 64 * ((width + 63) // 64) * ((height + 63) // 64)
        out = [self.empty(n, self.torch.float32) for _ in range(3)] if out is None else out
        arr = self._out3(out, self.torch.float32, n, "out")
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_dct64_image_dev(self.ctx, px, s, ph, width, height, int(method),
                                        1 if fused else 0, arr))
        return out

    def strategy_cost_sites(self, xyb, width, height, params, strategy, sites, qac,
                            meta_r=4.0, ytox=None, ytob=None, out=None):
        """Coefficient-model cost of `strategy` at each (bx, by) in `sites`
        ((n, 2) int tensor or array) with quantizers `qac` (n floats)."""
        t = self.torch
        sites = t.as_tensor(np.asarray(sites, np.int64)) if not t.is_tensor(sites) else sites
        packed = ((sites[:, 1].to(t.int64) << 16) | sites[:, 0].to(t.int64)).to(t.int32)
        packed = packed.to(self.device).contiguous()
        qac = t.as_tensor(qac, dtype=t.float32).to(self.device).contiguous()
        n = int(packed.shape[0])
        out = self.empty(n, t.float32) if out is None else out
        ytox, ytob = self._cmaps(width, height, ytox, ytob)
        nt = self._tiles(width, height)
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_strategy_cost_sites_dev(self.ctx, px, s, ph, width, height,
                                                self._buf(ytox, t.int8, nt, "ytox"),
                                                self._buf(ytob, t.int8, nt, "ytob"),
                                                ctypes.byref(params), int(strategy), n,
                                                self._buf(packed, t.int32, n, "sites"),
                                                self._buf(qac, t.float32, n, "qac"), float(meta_r),
                                                self._buf(out, t.float32, n, "out")))
        return out

    def chroma_hf_protection(self, xyb, width, height, qf, distance, fused=True):
        """apply_chroma_hf_protection on a device quant field, in place."""
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_chroma_hf_protection_dev(self.ctx, px, s, ph, width, height, float(distance),
                                                 1 if fused else 0,
                                                 self._buf(qf, self.torch.uint8,
                                                           self._blocks(width, height), "qf")))
        return qf

    def alpha_fringe_clear(self, alpha, width, height):
        """alpha_fringe::clear of a (H, W) uint8 device plane, in place."""
        check(lib().jxg_alpha_fringe_clear_dev(self.ctx, self._buf(alpha, self.torch.uint8,
                                                                   width * height, "alpha"),
                                               width, height))
        return alpha

    def fill_invisible(self, xyb, width, height, alpha):
        """fill_invisible of (3, padded_h, stride) planes under a (H, W) uint8
        alpha plane, in place (edge padding rewritten)."""
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_fill_invisible_dev(self.ctx, px, s, ph, width, height,
                                           self._buf(alpha, self.torch.uint8, width * height,
                                                     "alpha")))
        return xyb

    def alpha_block_factors(self, alpha, width, height, out=None):
        t = self.torch
        out = self.empty(((height + 7) // 8, (width + 7) // 8), t.float32) if out is None else out
        check(lib().jxg_alpha_block_factors_dev(self.ctx,
                                                self._buf(alpha, t.uint8, width * height, "alpha"),
                                                width, height,
                                                self._buf(out, t.float32, self._blocks(width, height),
                                                          "out")))
        return out

    def apply_alpha_factors(self, cells, factors, width, height, qf):
        """apply_alpha_factors on a device quant field, in place."""
        t, nb = self.torch, self._blocks(width, height)
        check(lib().jxg_apply_alpha_factors_dev(self.ctx, self._buf(cells, t.uint8, nb, "cells"),
                                                self._buf(factors, t.float32, nb, "factors"), width,
                                                height, self._buf(qf, t.uint8, nb, "qf")))
        return qf

    def alpha_plane(self, width, height):
        """The alpha plane the last front end call used, as a device tensor."""
        out = self.empty((height, width), self.torch.uint8)
        check(lib().jxg_alpha_plane(self.ctx, self._buf(out, self.torch.uint8, width * height, "out"),
                                    width * height, 0))
        return out

    def dct128_image(self, xyb, width, height, method=0, fused=True, out=None):
        """DCT128X128 of every 128x128 block (method 0: FP32 butterflies in
        the reference's order, 1: split-bf16 x3 on the matrix cores); xyb
        must reach whole 128-px blocks (see xyb_buffer128).  Returns 3 float
        tensors in the chunk layout with xblocks = 16 ceil(W/128)."""
        n = 16384 * ((width + 127) // 128) * ((height + 127) // 128)
        out = [self.empty(n, self.torch.float32) for _ in range(3)] if out is None else out
        arr = self._out3(out, self.torch.float32, n, "out")
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_dct128_image_dev(self.ctx, px, s, ph, width, height, int(method),
                                         1 if fused else 0, arr))
        return out

    def xyb_buffer128(self, width, height):
        """XYB planes padded to whole 128-px blocks (for dct128_image)."""
        return self.empty((3, (height + 127) // 128 * 128, (width + 127) // 128 * 128),
                          self.torch.float32)

    def quantize_status(self):
        out = (ctypes.c_uint32 * 2)()
        check(lib().jxg_quantize_status(self.ctx, out))
        return int(out[0]), int(out[1])

    def fill_cmap(self, xyb, width, height, params, ytox=None, ytob=None):
        """CfL maps (ytox, ytob) of (3, padded_h, stride) planes."""
        t = self.torch
        n = ((height + 63) // 64, (width + 63) // 64)
        ytox = self.empty(n, t.int8) if ytox is None else ytox
        ytob = self.empty(n, t.int8) if ytob is None else ytob
        nt = self._tiles(width, height)
        px, s, ph = self._planes(xyb, width, height)
        check(lib().jxg_fill_cmap_dev(self.ctx, px, s, ph, width, height, ctypes.byref(params),
                                      self._buf(ytox, t.int8, nt, "ytox"),
                                      self._buf(ytob, t.int8, nt, "ytob")))
        return ytox, ytob

    def run(self, rgb, width, height, params, outputs=None, qf=None, cells=None, ytox=None,
            ytob=None):
        """The whole front end on a device-resident (H, W, C) uint8 image.
        Returns levels / coeffs / dc plus the quant field, strategy cells and
        CfL maps; pass preallocated tensors to reuse them."""
        t = self.torch
        xb, yb = (width + 7) // 8, (height + 7) // 8
        n = ((height + 63) // 64, (width + 63) // 64)
        o = self.outputs(width, height) if outputs is None else outputs
        qf = self.empty((yb, xb), t.uint8) if qf is None else qf
        cells = self.empty((yb, xb), t.uint8) if cells is None else cells
        ytox = self.empty(n, t.int8) if ytox is None else ytox
        ytob = self.empty(n, t.int8) if ytob is None else ytob
        nb, nt = xb * yb, self._tiles(width, height)
        prgb, row_bytes, ch = self._image(rgb, width, height)
        lv, co, dc = self._outputs_arrays(o, width, height)
        check(lib().jxg_front_end_dev(self.ctx, prgb, row_bytes, ch, width, height,
                                      ctypes.byref(params), self._buf(ytox, t.int8, nt, "ytox"),
                                      self._buf(ytob, t.int8, nt, "ytob"),
                                      self._buf(qf, t.uint8, nb, "qf"),
                                      self._buf(cells, t.uint8, nb, "cells"), lv, co, dc))
        o = dict(o)
        o["qf"] = qf
        o["cells"] = cells
        o["ytox"] = ytox
        o["ytob"] = ytob
        return o


class Batch:
    """Many images through the front end at once (jxg_batch_*): one launch per
    stage for the whole batch, replayed as a HIP graph by launch().

    images: list of (rgb, params) with rgb a (H, W, 1|3) uint8 device tensor
    and params its FrameParams.  outputs(i) gives image i's results in the
    layout of FrontEnd.run.
    """

    def __init__(self, fe, images):
        from ._lib import BatchImage
        t = fe.torch
        self.fe = fe
        self.inputs = [rgb for rgb, _ in images]  # kept alive: the batch reads them
        self.outs = []
        arr = (BatchImage * len(images))()
        for i, (rgb, params) in enumerate(images):
            h, w, ch = int(rgb.shape[0]), int(rgb.shape[1]), int(rgb.shape[2])
            xb, yb = (w + 7) // 8, (h + 7) // 8
            tiles = ((h + 63) // 64, (w + 63) // 64)
            o = fe.outputs(w, h)
            o["qf"] = fe.empty((yb, xb), t.uint8)
            o["cells"] = fe.empty((yb, xb), t.uint8)
            o["ytox"] = fe.empty(tiles, t.int8)
            o["ytob"] = fe.empty(tiles, t.int8)
            self.outs.append(o)
            d = arr[i]
            prgb, row_bytes, _ = fe._image(rgb, w, h, f"images[{i}]")
            d.d_rgb = prgb.value
            d.row_bytes = row_bytes
            d.channels = ch
            d.width, d.height = w, h
            d.params = params
            d.d_ytox, d.d_ytob = o["ytox"].data_ptr(), o["ytob"].data_ptr()
            d.d_qf, d.d_cells = o["qf"].data_ptr(), o["cells"].data_ptr()
            for c in range(3):
                d.d_levels[c] = o["levels"][c].data_ptr()
                d.d_coeffs[c] = o["coeffs"][c].data_ptr()
                d.d_dc[c] = o["dc"][c].data_ptr()
        self.handle = ctypes.c_void_p()
        check(lib().jxg_batch_create(fe.ctx, len(images), arr, ctypes.byref(self.handle)))

    def run(self):
        """Enqueues the stage chain on the context's stream."""
        check(lib().jxg_batch_run(self.handle))

    def launch(self):
        """Runs the chain as a HIP graph (captured on first use)."""
        check(lib().jxg_batch_launch(self.handle))

    def status(self):
        out = (ctypes.c_uint32 * 2)()
        check(lib().jxg_batch_status(self.handle, out))
        return int(out[0]), int(out[1])

    def autotune(self, reps=3):
        buf = ctypes.create_string_buffer(4096)
        check(lib().jxg_batch_autotune(self.handle, int(reps), buf, len(buf)))
        return buf.value.decode()

    def config(self):
        out = (ctypes.c_int * 3)()
        check(lib().jxg_batch_config(self.handle, out))
        return dict(aq_strip=out[0], cm_rows=out[1], xt_grid=out[2])

    def outputs(self, i):
        return self.outs[i]

    def close(self):
        if self.handle:
            lib().jxg_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QfTuner:
    """Gradient tuning of the quant field (jxg_qf_tuner_*) for one image:
    QfTuner(fe, xyb, width, height, cells, params, lam=0.01, seed=1).
    loss_grad(theta, step) -> (loss, grad); run(qf, steps, lr) -> (integer
    field, real field, loss trace)."""

    def __init__(self, fe, xyb, width, height, cells, params, lam=0.01, seed=1, max_steps=64):
        self.fe = fe
        self.keep = (xyb, cells)  # the tuner reads them at every step
        self.width, self.height = width, height
        self.handle = ctypes.c_void_p()
        px, s, ph = fe._planes(xyb, width, height)
        pc = fe._buf(cells, fe.torch.uint8, fe._blocks(width, height), "cells")
        check(lib().jxg_qf_tuner_create(fe.ctx, px, s, ph, width,
                                        height, pc, ctypes.byref(params), float(lam),
                                        int(seed), int(max_steps), ctypes.byref(self.handle)))

    def excluded(self):
        n = ctypes.c_uint32()
        check(lib().jxg_qf_tuner_excluded(self.handle, ctypes.byref(n)))
        return int(n.value)

    def loss_grad(self, theta, step=0):
        t = self.fe.torch
        theta = theta.to(t.float32).contiguous()
        grad = t.empty_like(theta)
        loss = ctypes.c_float()
        nb = self.fe._blocks(self.width, self.height)
        check(lib().jxg_qf_tuner_loss_grad(self.handle, self.fe._buf(theta, t.float32, nb, "theta"),
                                           int(step), self.fe._buf(grad, t.float32, nb, "grad"),
                                           ctypes.byref(loss)))
        return float(loss.value), grad

    def run(self, qf, steps=20, lr=0.05):
        t = self.fe.torch
        out = t.empty_like(qf)
        theta = t.empty(qf.shape, dtype=t.float32, device=qf.device)
        trace = (ctypes.c_float * (steps + 1))()
        nb = self.fe._blocks(self.width, self.height)
        check(lib().jxg_qf_tuner_run(self.handle, self.fe._buf(qf, t.uint8, nb, "qf"), int(steps),
                                     float(lr), self.fe._buf(out, t.uint8, nb, "out"),
                                     self.fe._buf(theta, t.float32, nb, "theta"), trace))
        return out, theta, [float(v) for v in trace]

    def close(self):
        if self.handle:
            lib().jxg_qf_tuner_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LumaProxy:
    """The EPF ranking proxy on the GPU (LumaProxy, src/epf_proxy.rs:151):
    LumaProxy(front_end, y) with y a (H, W) float32 device tensor (or a 2-D
    view with unit column stride); compare(candidate) -> (score, terms)."""

    def __init__(self, fe, y):
        self.fe = fe
        h, w = y.shape
        self.shape = (h, w)
        self.ptr = ctypes.c_void_p()
        check(lib().jxg_epf_proxy_new(fe.ctx, self._rows(y), y.stride(0), w, h,
                                      ctypes.byref(self.ptr)))

    def _rows(self, y):
        """(H, W) float32 rows with unit column stride on the context's device."""
        t = self.fe.torch
        if not isinstance(y, t.Tensor) or y.dim() != 2:
            raise JxgError(-1, "y: expected an (H, W) tensor")
        if not y.is_cuda or y.device != self.fe.device or y.dtype != t.float32:
            raise JxgError(-1, f"y: expected float32 on {self.fe.device}, got {y.dtype} on {y.device}")
        if y.stride(1) != 1 or y.stride(0) < y.shape[1]:
            raise JxgError(-1, f"y: rows must have unit column stride (strides {y.stride()})")
        return ctypes.c_void_p(y.data_ptr())

    def compare(self, y):
        if tuple(y.shape) != self.shape:
            raise JxgError(-1, f"candidate shape {tuple(y.shape)} != {self.shape}")
        score = ctypes.c_float()
        terms = np.zeros(6, np.float32)
        check(lib().jxg_epf_proxy_compare(self.ptr, self._rows(y), y.stride(0), ctypes.byref(score),
                                          terms.ctypes.data))
        return float(score.value), terms

    def close(self):
        if self.ptr:
            lib().jxg_epf_proxy_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def encode_front_end(rgb, config: EncodeConfig = EncodeConfig(), device=0,
                     tables=None) -> FrontEndOutput:
    """Host-buffer front end of one 8-bit image on `device`: (H, W, 3|4) RGB(A)
    or (H, W[, 1|2]) gray (+ alpha), the gray path of encode_image_gray.

    The drop-in counterpart of the front-end half of jixel's encode_image /
    encode_image_with_alpha (src/encode_image.rs:1610,3482): the returned
    buffers are what the CPU back end (RDOQ, tokenization, entropy coding)
    consumes.
    """
    rgb = np.ascontiguousarray(rgb, np.uint8)
    if rgb.ndim == 2:
        rgb = rgb[:, :, None]
    if rgb.ndim != 3 or rgb.shape[2] not in (1, 2, 3, 4):
        raise JxgError(-1, f"expected (H, W[, 1|2|3|4]) uint8, got {rgb.shape}")
    h, w, ch = rgb.shape
    if h == 0 or w == 0:
        raise JxgError(-2, "empty image")
    fe = FrontEnd(device, tables=tables)
    try:
        p = frame_params(config)
        xb, yb = (w + 7) // 8, (h + 7) // 8
        tiles = ((h + 63) // 64, (w + 63) // 64)
        qf = np.zeros((yb, xb), np.uint8)
        cells = np.zeros((yb, xb), np.uint8)
        ytox = np.zeros(tiles, np.int8)
        ytob = np.zeros(tiles, np.int8)
        levels = [np.zeros(yb * xb * 64, np.int32) for _ in range(3)]
        coeffs = [np.zeros(yb * xb * 64, np.float32) for _ in range(3)]
        dc = [np.zeros(yb * xb, np.int16) for _ in range(3)]
        arr = lambda xs: (ctypes.c_void_p * 3)(*[x.ctypes.data for x in xs])
        check(lib().jxg_front_end(fe.ctx, rgb.ctypes.data, w * ch, ch, w, h, ctypes.byref(p),
                                  ytox.ctypes.data, ytob.ctypes.data, qf.ctypes.data,
                                  cells.ctypes.data, arr(levels), arr(coeffs), arr(dc)))
        alpha = None
        if ch in (2, 4):
            alpha = np.zeros((h, w), np.uint8)
            check(lib().jxg_alpha_plane(fe.ctx, alpha.ctypes.data, alpha.size, 1))
        return FrontEndOutput(qf=qf, cells=cells, ytox=ytox, ytob=ytob, levels=levels,
                              coeffs=coeffs, dc=dc, alpha=alpha)
    finally:
        fe.close()
