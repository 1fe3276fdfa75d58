"""GPU parity of the whole-plane filters against the oracle, through the C
ABI.  The gaborish kernel follows the reference's operation order, so the
filtered planes must match bit for bit, and the padding past the image must
be the edge replication of the filtered image."""
import numpy as np
import pytest

from test_gpu_parity import J, fe, torch, padded, same_bits, first_mismatch, log  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("w,h", [(1, 1), (3, 2), (67, 45), (256, 130), (130, 300)])
def test_gaborish_bit_exact(J, fe, torch, oracle, w, h, fused):
    log(f"gaborish {w}x{h} fused={fused}")
    rng = np.random.default_rng(w + 7 * h)
    xyb = rng.uniform(-0.05, 0.9, size=(3, h, w)).astype(np.float32)
    oracle.set_fused(fused)
    want = oracle.gaborish_inverse(xyb)
    oracle.set_fused(True)
    d = torch.from_numpy(padded(J, xyb)).cuda()
    sep = fe.gaborish_inverse(d, w, h, fused=fused, out=torch.empty_like(d)).cpu().numpy()
    fe.gaborish_inverse(d, w, h, fused=fused)  # in place
    got = d.cpu().numpy()
    m = first_mismatch(got[:, :h, :w], want)
    assert m is None, m
    assert same_bits(got, padded(J, want)), "padding is not the edge replication"
    assert same_bits(sep, got), "out-of-place and in-place results differ"


def test_gaborish_rejects_negative_mul(J, fe, torch):
    d = torch.zeros((3, 64, 64), dtype=torch.float32, device="cuda")
    with pytest.raises(J.JxgError) as e:
        fe.gaborish_inverse(d, 8, 8, mul=-1.0)
    assert e.value.code == -1
