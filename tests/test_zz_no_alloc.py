"""The no-allocation entry point jxg_fill_quant_field_noalloc_dev, called
through the C ABI, against jxg_fill_quant_field_dev on the same inputs: two
calls in a row, the second on a smaller image (the static scratch keeps the
first call's contents), bit for bit; plus the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(J, torch, lib, fe, entry, rgb, distance, fused):
    h, w, _ = rgb.shape
    xyb = fe.xyb(torch.from_numpy(rgb).cuda(), w, h, fused=fused)
    s, ph = J.plane_geometry(w, h)
    qf = torch.full(((h + 7) // 8, (w + 7) // 8), 77, dtype=torch.uint8, device="cuda")
    p = J.frame_params(J.EncodeConfig(distance=distance, fused=fused))
    rc = entry(fe.ctx, ctypes.c_void_p(xyb.data_ptr()), s, ph, w, h, ctypes.c_float(distance),
               ctypes.c_float(1.0 / p.scale), 1 if fused else 0, ctypes.c_void_p(qf.data_ptr()))
    assert rc == 0, lib.jxg_last_error()
    torch.cuda.synchronize()
    return qf.cpu().numpy()


def test_noalloc_quant_field_matches_the_allocating_entry_point():
    import torch

    import jixel_amd as J
    from jixel_amd import _lib, synth
    lib = _lib.lib()
    fe = J.FrontEnd(0)
    try:
        cases = [(synth.photo_like(640, 480, 21), 1.0, True),
                 (synth.gradient_sines(200, 136, 5), 0.5, False)]
        for rgb, d, fused in cases:
            print("noalloc", rgb.shape, d, fused, flush=True)
            want = _run(J, torch, lib, fe, lib.jxg_fill_quant_field_dev, rgb, d, fused)
            got = _run(J, torch, lib, fe, lib.jxg_fill_quant_field_noalloc_dev, rgb, d, fused)
            assert np.array_equal(got, want)
            assert len(np.unique(want)) > 1  # a modulated field, not the flat one
        # larger than the static pre-map: rejected before any launch
        w = h = 16384
        s, ph = J.plane_geometry(w, h)
        dummy = torch.zeros(16, dtype=torch.float32, device="cuda")
        rc = lib.jxg_fill_quant_field_noalloc_dev(fe.ctx, ctypes.c_void_p(dummy.data_ptr()), s, ph,
                                                   w, h, ctypes.c_float(1.0), ctypes.c_float(10.0),
                                                   1, ctypes.c_void_p(dummy.data_ptr()))
        assert rc == -3, rc  # JXG_ERR_TOO_LARGE
        assert b"no-allocation" in lib.jxg_last_error()
    finally:
        fe.close()
