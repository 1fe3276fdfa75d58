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


def test_frame_params_carry_the_spec_xyb_matrix():
    p = J.frame_params(J.EncodeConfig())
    assert np.array_equal(np.array(p.xyb_fwd[:], np.float32), J.XYB_SPEC_FWD)
    assert p.xyb_order == 0 and p.arbiter == J.ARBITER_SSIMULACRA2
    out = np.zeros(9, np.float32)
    _lib.lib().jxg_spec_xyb_matrix(out.ctypes.data)
    assert np.array_equal(out, J.XYB_SPEC_FWD)
    q = J.frame_params(J.EncodeConfig(xyb_order=1).with_arbiter(J.ARBITER_CVVDP))
    assert q.xyb_order == 1 and q.arbiter == 1


def test_device_pointer_check_rejects_host_tensors():
    """The wrappers' pointer check (the C-ABI checks only for null): a host
    tensor, a wrong dtype, a strided view and a short buffer are refused
    before any launch."""
    import torch
    import jixel_amd as J
    from jixel_amd import _dev_ptr
    dev = torch.device("cuda", 0)
    t = torch.zeros(16, dtype=torch.uint8)
    for bad, text in ((t, "expected a tensor on"), ([1, 2], "expected a torch tensor")):
        with pytest.raises(J.JxgError) as e:
            _dev_ptr(bad, dev, torch.uint8, 16, "buf")
        assert e.value.code == -1 and text in str(e.value)
