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


@pytest.mark.parametrize("w,h", [(4, 4), (9, 7), (257, 129), (640, 360)])
def test_epf_proxy_matches_oracle(J, fe, torch, oracle, w, h):
    """Per-pixel terms follow the reference's f32 order; the pooled sums are
    f64 on the GPU: equal to the oracle's f64 pooling within 1e-6 relative,
    and to the reference's f32 pixel-order pooling within 1e-4."""
    log(f"epf proxy {w}x{h}")
    rng = np.random.default_rng(w * 7 + h)
    src = rng.uniform(0.05, 0.85, (h, w)).astype(np.float32)
    cand = (src + rng.normal(0, 0.015, src.shape)).astype(np.float32)
    d_src = torch.from_numpy(src).cuda()
    proxy = J.LumaProxy(fe, d_src)
    try:
        score, terms = proxy.compare(torch.from_numpy(cand).cuda())
        ident, iterms = proxy.compare(d_src)
    finally:
        proxy.close()
    s64, t64 = oracle.epf_compare(src, cand, wide=True)
    s32, t32 = oracle.epf_compare(src, cand, wide=False)
    assert np.allclose(terms, t64, rtol=1e-6, atol=1e-12), (terms, t64)
    assert abs(score - s64) <= 1e-6 * abs(s64)
    assert np.allclose(terms, t32, rtol=1e-4, atol=1e-9) and abs(score - s32) <= 1e-4 * abs(s32)
    assert ident == 0.0 and not iterms.any()


def test_epf_proxy_strided_candidate_and_errors(J, fe, torch, oracle):
    rng = np.random.default_rng(3)
    src = rng.uniform(0.1, 0.9, (64, 80)).astype(np.float32)
    big = np.zeros((64, 96), np.float32)
    big[:, :80] = src + 0.01
    d_src = torch.from_numpy(src).cuda()
    proxy = J.LumaProxy(fe, d_src)
    try:
        s_view, _ = proxy.compare(torch.from_numpy(big).cuda()[:, :80])
        s_dense, _ = proxy.compare(torch.from_numpy(np.ascontiguousarray(big[:, :80])).cuda())
    finally:
        proxy.close()
    assert s_view == s_dense
    with pytest.raises(J.JxgError) as e:
        J.LumaProxy(fe, torch.zeros((3, 80), dtype=torch.float32, device="cuda"))
    assert e.value.code == -1
