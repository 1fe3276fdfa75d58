"""The GPU strategy search on the reference's Fast-scope known answer
(src/ac_strategy/selection.rs:4107-4150): a flat 32x32 region at distance
3, global scale 1, quant field 8 merges to DCT32X32 and offers no other
shape, and the cells equal the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h", [(32, 32), (96, 64)])
def test_flat_region_merges_to_dct32x32(oracle, fused, w, h):
    import torch
    import jixel_amd as J
    fe = J.FrontEnd(0)
    try:
        oracle.set_fused(fused)
        p = J.frame_params(J.EncodeConfig().with_distance(3.0))
        p.scale = 1.0
        p.x_qm_scale = 2
        p.fused = 1 if fused else 0
        xyb = fe.xyb_buffer(w, h).zero_()
        qf = torch.full((h // 8, w // 8), 8, dtype=torch.uint8, device="cuda")
        cells = fe.select_strategy(xyb, w, h, qf, p).cpu().numpy()
        S = J.STRATEGY
        want, _ = oracle.select_strategy(np.zeros((3, h, w), np.float32),
                                         np.full((h // 8, w // 8), 8, np.uint8),
                                         np.zeros(((h + 63) // 64, (w + 63) // 64), np.int8),
                                         np.zeros(((h + 63) // 64, (w + 63) // 64), np.int8),
                                         oracle.tables_for(p), p)
        assert np.array_equal(cells, want)
        for by in range(0, h // 8, 4):
            for bx in range(0, w // 8, 4):
                assert cells[by, bx] >> 1 == S["DCT32X32"] and cells[by, bx] & 1, (bx, by)
        firsts = {int(c >> 1) for c in cells.reshape(-1) if c & 1}
        assert firsts == {S["DCT32X32"]}
    finally:
        fe.close()
