"""Seeded synthetic images for the front end's workloads.

* photo_like(w, h, seed): 1/f-spectrum noise (random phases, amplitude
  proportional to 1/|k|, mean 128) + 40 flat axis-aligned rectangles with
  sharp edges + Gaussian noise sigma 3, clamped to u8 (the 2048x2048 q75
  workload and the batch/gradient workloads).
* gradient_sines(w, h, seed): diagonal gradient 0..255 + three sinusoids of
  amplitude 20 and period 16-64 px in random directions + Gaussian noise
  sigma 4 (the 256x256 q90 workload).
* smooth_rgba(w, h, seed): 1/f noise box-smoothed over 32-128 px regions with
  an alpha of 0/255 inside 20 soft-edged ellipses (the 7680x4320 RGBA case);
  smooth_rgba_device is the same recipe generated on the GPU.

PRNG: numpy PCG64 seeded with the workload's seed; all float math in f64,
rounded to u8 at the end.
"""
import hashlib

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _one_over_f(rng, w, h, channels=3, contrast=48.0):
    out = np.empty((h, w, channels))
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    k = np.sqrt(fx * fx + fy * fy)
    k[0, 0] = 1.0
    amp = 1.0 / k
    amp[0, 0] = 0.0
    for c in range(channels):
        phase = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
        spec = amp * np.exp(1j * phase)
        img = np.fft.irfft2(spec, s=(h, w))
        img /= img.std() + 1e-12
        out[..., c] = img * contrast
    # correlate the channels a little, as photos are
    mix = out.mean(axis=2, keepdims=True)
    return 128.0 + 0.6 * mix + 0.4 * out


def photo_like(w, h, seed):
    rng = _rng(seed)
    img = _one_over_f(rng, w, h)
    for _ in range(40):
        rw = int(rng.integers(max(2, w // 64), max(3, w // 4)))
        rh = int(rng.integers(max(2, h // 64), max(3, h // 4)))
        x0 = int(rng.integers(0, max(1, w - rw)))
        y0 = int(rng.integers(0, max(1, h - rh)))
        img[y0:y0 + rh, x0:x0 + rw, :] = rng.uniform(0, 255, size=3)
    img += rng.normal(0.0, 3.0, size=img.shape)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def one_over_f_field(side, seed, device):
    """A 3-channel 1/f-spectrum field of side x side on the device (torch;
    one FFT size), from which batch images take crops."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    fy = torch.fft.fftfreq(side, device=device, dtype=torch.float64)[:, None]
    fx = torch.fft.rfftfreq(side, device=device, dtype=torch.float64)[None, :]
    k = torch.sqrt(fx * fx + fy * fy)
    k[0, 0] = 1.0
    amp = 1.0 / k
    amp[0, 0] = 0.0
    chans = []
    for _ in range(3):
        phase = torch.rand(amp.shape, generator=g, device=device, dtype=torch.float64) * (2.0 * np.pi)
        chans.append(torch.fft.irfft2(torch.polar(amp, phase), s=(side, side)))
    return torch.stack(chans, dim=2)


def photo_like_device(w, h, seed, device, field):
    """photo_like's recipe on the device: a crop of a shared 1/f field at a
    seeded offset, normalized per channel like photo_like, with 40 flat
    rectangles and Gaussian noise sigma 3 (torch random stream: the same
    distribution as photo_like, different pixels).  Used for the 512-image
    batch workload, whose host generation would take minutes."""
    import torch
    rng = _rng(seed)
    side = field.shape[0]
    ox, oy = int(rng.integers(0, side - w + 1)), int(rng.integers(0, side - h + 1))
    crop = field[oy:oy + h, ox:ox + w, :]
    out = crop / (crop.std(dim=(0, 1), keepdim=True) + 1e-12) * 48.0
    img = 128.0 + 0.6 * out.mean(dim=2, keepdim=True) + 0.4 * out
    rects = torch.tensor(rng.uniform(0, 255, size=(40, 3)), device=device)
    for r in range(40):
        rw = int(rng.integers(max(2, w // 64), max(3, w // 4)))
        rh = int(rng.integers(max(2, h // 64), max(3, h // 4)))
        x0 = int(rng.integers(0, max(1, w - rw)))
        y0 = int(rng.integers(0, max(1, h - rh)))
        img[y0:y0 + rh, x0:x0 + rw, :] = rects[r]
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    img += torch.randn(img.shape, generator=g, device=device, dtype=torch.float64) * 3.0
    return img.round().clamp(0, 255).to(torch.uint8).contiguous()


def batch_sizes(n, seed):
    """(width, height, quality) of the batch workload: sides uniform in
    [64, 2048] and not multiples of 8, quality uniform in [50, 95]."""
    rng = _rng(seed)
    out = []
    for _ in range(n):
        w, h = (int(v) for v in rng.integers(64, 2049, size=2))
        w = w - 1 if w % 8 == 0 else w
        h = h - 1 if h % 8 == 0 else h
        out.append((w, h, float(rng.uniform(50.0, 95.0))))
    return out


def gradient_sines(w, h, seed):
    rng = _rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    base = (x + y) * 255.0 / max(1, (w + h - 2))
    img = np.repeat(base[..., None], 3, axis=2)
    for _ in range(3):
        period = rng.uniform(16.0, 64.0)
        theta = rng.uniform(0.0, np.pi)
        phase = rng.uniform(0.0, 2.0 * np.pi)
        wave = 20.0 * np.sin(2.0 * np.pi * (x * np.cos(theta) + y * np.sin(theta)) / period + phase)
        img += wave[..., None] * rng.uniform(0.5, 1.0, size=3)
    img += rng.normal(0.0, 4.0, size=img.shape)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def _box_smooth(img, r):
    # separable running mean with edge replication
    out = img
    for axis in (0, 1):
        pad = [(0, 0)] * out.ndim
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="edge")
        c = np.cumsum(p, axis=axis)
        c = np.concatenate([np.zeros_like(np.take(c, [0], axis=axis)), c], axis=axis)
        n = out.shape[axis]
        hi = np.take(c, np.arange(2 * r + 1, 2 * r + 1 + n), axis=axis)
        lo = np.take(c, np.arange(0, n), axis=axis)
        out = (hi - lo) / (2 * r + 1)
    return out


def smooth_rgba(w, h, seed):
    rng = _rng(seed)
    rgb = _one_over_f(rng, w, h, contrast=60.0) - 128.0
    r = int(rng.integers(16, 65))
    rgb = 128.0 + _box_smooth(rgb, r) * 3.0
    rgb += rng.normal(0.0, 2.0, size=rgb.shape)
    alpha = np.zeros((h, w))
    for _ in range(20):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        ax, ay = rng.uniform(w / 40, w / 6), rng.uniform(h / 40, h / 6)
        # outside the bounding box d > 1 and the falloff is 0, so only the box is visited
        x0, x1 = max(0, int(np.floor(cx - ax))), min(w, int(np.ceil(cx + ax)) + 1)
        y0, y1 = max(0, int(np.floor(cy - ay))), min(h, int(np.ceil(cy + ay)) + 1)
        if x0 >= x1 or y0 >= y1:
            continue
        y, x = np.mgrid[y0:y1, x0:x1].astype(np.float64)
        d = np.sqrt(((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2)
        edge = np.clip((1.0 - d) * min(ax, ay) / 4.0, 0.0, 1.0)  # 4 px falloff
        alpha[y0:y1, x0:x1] = np.maximum(alpha[y0:y1, x0:x1], edge)
    out = np.empty((h, w, 4))
    out[..., :3] = rgb
    out[..., 3] = alpha * 255.0
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def smooth_rgba_device(w, h, seed, device):
    """smooth_rgba's recipe on the device (torch, f64): 1/f noise of contrast
    60 box-smoothed with a seeded radius in [16, 64] (a 33..129 px window),
    x3 around 128, Gaussian noise sigma 2, and an alpha of 0 / 255 inside 20
    seeded ellipses with a 4 px linear falloff.  The radius and ellipses are
    drawn from a PCG64 stream with the workload's seed, the noise fields from
    torch's generator: the same distribution as smooth_rgba, other pixels.
    Used for the 7680x4320 benchmark image, whose host generation takes half
    a minute."""
    import torch
    rng = _rng(seed)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    fy = torch.fft.fftfreq(h, device=device, dtype=torch.float64)[:, None]
    fx = torch.fft.rfftfreq(w, device=device, dtype=torch.float64)[None, :]
    k = torch.sqrt(fx * fx + fy * fy)
    k[0, 0] = 1.0
    amp = 1.0 / k
    amp[0, 0] = 0.0
    chans = []
    for _ in range(3):
        phase = torch.rand(amp.shape, generator=g, device=device, dtype=torch.float64) * (2.0 * np.pi)
        img = torch.fft.irfft2(torch.polar(amp, phase), s=(h, w))
        chans.append(img / (img.std() + 1e-12) * 60.0)
    out = torch.stack(chans, dim=2)
    rgb = 0.6 * out.mean(dim=2, keepdim=True) + 0.4 * out
    r = int(rng.integers(16, 65))
    for axis in (0, 1):  # separable running mean with edge replication
        n = rgb.shape[axis]
        idx = torch.clamp(torch.arange(-r, n + r, device=device), 0, n - 1)
        p = torch.index_select(rgb, axis, idx)
        c = torch.cumsum(p, dim=axis)
        c = torch.cat([torch.zeros_like(torch.narrow(c, axis, 0, 1)), c], dim=axis)
        rgb = (torch.narrow(c, axis, 2 * r + 1, n) - torch.narrow(c, axis, 0, n)) / (2 * r + 1)
    rgb = 128.0 + rgb * 3.0
    rgb += torch.randn(rgb.shape, generator=g, device=device, dtype=torch.float64) * 2.0
    alpha = torch.zeros((h, w), device=device, dtype=torch.float64)
    for _ in range(20):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        ax, ay = rng.uniform(w / 40, w / 6), rng.uniform(h / 40, h / 6)
        x0, x1 = max(0, int(np.floor(cx - ax))), min(w, int(np.ceil(cx + ax)) + 1)
        y0, y1 = max(0, int(np.floor(cy - ay))), min(h, int(np.ceil(cy + ay)) + 1)
        if x0 >= x1 or y0 >= y1:
            continue
        y = torch.arange(y0, y1, device=device, dtype=torch.float64)[:, None]
        x = torch.arange(x0, x1, device=device, dtype=torch.float64)[None, :]
        d = torch.sqrt(((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2)
        edge = torch.clamp((1.0 - d) * min(ax, ay) / 4.0, 0.0, 1.0)
        alpha[y0:y1, x0:x1] = torch.maximum(alpha[y0:y1, x0:x1], edge)
    img = torch.cat([rgb, (alpha * 255.0)[..., None]], dim=2)
    return img.round().clamp(0, 255).to(torch.uint8).contiguous()


def sha256(img):
    return hashlib.sha256(np.ascontiguousarray(img).tobytes()).hexdigest()


WORKLOADS = {
    # name: (generator, width, height, seed, quality)
    "gradient_256_q90": (gradient_sines, 256, 256, 1, 90.0),
    "photo_2048_q75": (photo_like, 2048, 2048, 2, 75.0),
    "smooth_rgba_7680x4320_q75": (smooth_rgba, 7680, 4320, 3, 75.0),
}


def workload(name):
    gen, w, h, seed, q = WORKLOADS[name]
    return gen(w, h, seed), q
