"""Closed-form pins of the oracle's candidate cost for every strategy up to
64x64: a flat image has only DC energy, which the cost skips, so the cost is
the metadata charge alone (lambda * meta_r); a single nonzero AC level gives
the model's rate terms."""
import numpy as np
import pytest

NAMES = ["DCT", "DCT16X8", "DCT8X16", "DCT16X16", "DCT32X32", "DCT32X16", "DCT16X32", "DCT32X8",
         "DCT8X32", "DCT64X64", "DCT64X32", "DCT32X64", "IDENTITY", "DCT2X2", "DCT4X4", "DCT4X8",
         "DCT8X4", "AFV0", "AFV1", "AFV2", "AFV3"]
RD_LAMBDA = np.float32(0.08086717)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("meta_r", [4.0, 0.0])
def test_flat_image_costs_only_metadata(oracle, name, meta_r):
    import jixel_amd as J
    xyb = np.empty((3, 64, 64), np.float32)
    xyb[0], xyb[1], xyb[2] = 0.01, 0.4, 0.38
    p = J.frame_params(J.EncodeConfig().with_distance(1.5))
    s = J.STRATEGY[name]
    cost = oracle.strategy_cost(xyb, s, 0, 0, float(p.scale * 40), [0.0, 0.0, 1.0],
                                oracle.tables_for(p), p, meta_r=meta_r)
    if name.startswith("AFV"):
        # the AFV basis leaves f32 rounding residue in its AC slots: no level,
        # a squared error of order 1e-15
        assert 0.0 <= cost - RD_LAMBDA * np.float32(meta_r) < 1e-6
    else:
        assert cost == RD_LAMBDA * np.float32(meta_r)
