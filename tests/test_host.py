"""Host side of the product (no GPU needed): the C ABI loads and exports every
symbol include/jxg.h declares, and its parameter helpers agree with the
reference's known answers and with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import jixel_amd as J
from jixel_amd import _lib


def header_functions():
    with open(os.path.join(ROOT, "include", "jxg.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(jxg_[a-z_0-9]+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(_lib.EXPORTED) <= set(header_functions())


def test_quality_to_distance_matches_oracle(oracle):
    for q in [0, 10, 29.5, 30, 50, 75, 80.5, 90, 95, 99, 99.5, 100, 120, -5]:
        assert J.distance_from_quality(q) == oracle.scalar("distance_from_quality", q)
    with pytest.raises(ValueError):
        J.distance_from_quality(float("nan"))


@pytest.mark.parametrize("distance", [0.1, 0.278, 0.5, 1.0, 1.3, 2.0, 2.35, 4.6, 6.0, 9.5, 12.0, 25.0])
@pytest.mark.parametrize("fused", [True, False])
def test_frame_params_match_oracle(oracle, distance, fused):
    d = float(np.float32(distance))
    oracle.set_fused(fused)
    try:
        want = oracle.distance_params(d)
    finally:
        oracle.set_fused(True)
    p = J.frame_params(J.EncodeConfig(distance=d, fused=fused))
    assert p.global_scale == want["global_scale"]
    assert p.scale == want["scale"] and p.scale_dc == want["scale_dc"]
    assert p.quant_dc == want["quant_dc"] and p.x_qm_scale == want["x_qm_scale"]
    assert p.epf_iters == want["epf_iters"]
    assert list(p.dc_step) == list(want["dc_step"])
    assert p.cfl_color_factor == 84.0 and p.cfl_base_b == 1.0 and p.b_qm_scale == 2


def test_frame_params_rejects_bad_distance():
    for d in [0.0, -1.0, float("inf"), float("nan")]:
        p = J.FrameParams()
        assert _lib.lib().jxg_frame_params_init(d, 2, 1, ctypes.byref(p)) == -4


def test_srgb_lut_matches_oracle(oracle):
    assert np.array_equal(J.srgb_lut(), oracle.srgb_lut())


def test_table_blob_layout():
    sizes = [J.qt_size(k) for k in range(13)]
    assert sizes == [64, 128, 256, 1024, 512, 256, 4096, 2048, 64, 64, 64, 64, 64]
    assert J.qt_offset(13) == J.QT_BLOB_FLOATS == 6 * sum(sizes)
    assert [_lib.lib().jxg_qt_kind_of(s) for s in range(24)][:5] == [0, 1, 1, 2, 3]


def test_default_tables_match_reference_library_tables(golden):
    """The band construction reproduces the reference's stored DCT8 and 16x8
    tables to the ~4e-6 the reference quotes between them
    (src/quant_weights.rs:1996-1999)."""
    blob = J.default_quant_tables()
    for kind, name, n in [(0, "DEQUANT_MATRIX_8X8", 64), (1, "DEQUANT_MATRIX_16X8", 128)]:
        off = J.qt_offset(kind)
        for c in range(3):
            got = blob[off + c * n: off + (c + 1) * n]
            want = np.array(golden[name][c], np.float32)
            assert np.allclose(got, want, rtol=2e-5, atol=0), (name, c)
            inv = blob[off + (3 + c) * n: off + (4 + c) * n]
            assert inv[0] == 0.0
            assert np.abs(got[1:] * inv[1:] - 1.0).max() <= 4 * np.finfo(np.float32).eps


def test_dct16_and_32_tables_dc_weight():
    # dct16x16_matrix_dc_matches_libjxl_polynomial (src/quant_weights.rs:3477)
    blob = J.default_quant_tables()
    off = J.qt_offset(2)
    for c, band0 in enumerate([8996.87257, 3191.48366, 1157.50408]):
        assert abs(blob[off + c * 256] - 1.0 / band0) < 1e-7
        m = blob[off + c * 256: off + (c + 1) * 256]
        assert (m > 0).all() and m[255] > m[0]


def test_plane_geometry():
    assert J.plane_geometry(1, 1) == (64, 64)
    assert J.plane_geometry(2048, 2048) == (2048, 2048)
    assert J.plane_geometry(7680, 4320) == (7680, 4352)
    assert J.plane_geometry(65, 129) == (128, 192)


def test_config_mirror():
    c = J.EncodeConfig()
    assert c.distance == 1.0 and c.speed == J.Speed.FAST
    assert abs(c.with_quality(75).distance - 2.35) < 1e-6
    with pytest.raises(J.JxgError):
        c.with_distance(0.0)
    assert c.with_speed(4).speed == J.Speed.MEDIUMSLOW


def test_no_device_is_a_loud_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(J.JxgError):
        J.FrontEnd(0)
