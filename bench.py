"""VarDCT front-end throughput on one MI355X.

A step is one pass of the front end (u8 sRGB(A) -> XYB, alpha fringe
clearing and fill_invisible -> adaptive quant field -> CfL maps -> AC-strategy
search -> alpha block factors -> forward DCT + quantization) over one
7680x4320 RGBA8 synthetic image (1/f noise smoothed over 33-129 px regions,
alpha 0 / 255 inside 20 soft-edged ellipses) at quality 75, at the
reference's default speed, with the input already resident in HBM and the
outputs left there.  Prints one JSON line: whole-job megapixels per second,
per-kernel device times measured inside the timed region, the dominant
kernel's roofline position, the pinned-host pipelined rate (PCIe included,
never the headline), the 2048x2048 RGB8 q75 photo-like workload, the batch
and quant-field tuning workloads, and the CPU oracle on the host cores of the
same machine as a reported baseline.

usage: python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, GB/s
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32, TFLOP/s

# Algorithmic HBM bytes per pixel of each profiled kernel (reads + writes of
# the buffers it must touch once), for RGBA input.
KERNEL_BYTES_PER_PX = {
    "xyb": 4.0 + 12.0,           # u8 RGBA in, 3 f32 planes out
    "alpha_prep": 4.0 + 1.0,     # RGBA in for its alpha bytes, the alpha plane out
    "fill": 12.0 + 12.0 + 1.0,   # fill_invisible: 3 planes read and written, alpha in
    "cmap": 12.0,                # 3 planes in, two i8 per 64x64 tile out
    "aq_pre": 4.0 + 1.0 / 4,     # Y plane in, the 1/16-size pre-map out
    "aq_block": 12.0 + 1.0 / 64, # 3 planes in, u8 field out
    "select": 12.0 + 2.0 / 64,   # 3 planes in, cells + field out
    "alpha_qf": 1.0 + 2.0 / 64,  # alpha in, field rewritten
    "quantize": 12.0 + 24.0 + 6.0 / 64 + 2.0 / 64,  # planes in, i32 + f32 out, DC, qf, cells
    "quantize_xt": 0.0,          # the tile-crossing queue (its bytes are a sliver of quantize's)
    "retile": 12.0 + 1.0 / 64,   # 3 planes in (the piece costs), field and cells rewritten
    "refine": 12.0 + 2.0 / 64,   # 3 planes in, cells in, field rewritten
    "rerank": 12.0 + 2.0 / 64,   # 3 planes in, field in, cells rewritten (inside "retile")
}
# FP32 work of the strategy search per pixel: the DCT8 / 16x16 / 32x32
# candidates' butterflies (4 log2 N flops per pixel per channel for an N x N
# transform: 12 + 16 + 20) x 3 channels, plus about 10 flops per coefficient
# for the quantize / SSE / rate terms of 3 candidates x 3 channels.
SELECT_FLOPS_PER_PX = 3 * (12 + 16 + 20) + 3 * 3 * 10
# FP32 work of the re-tiling's piece costs per pixel: at every block a DCT8,
# a 16x8, an 8x16 and a 16x16 (2 log2 N flops per pixel per axis: 768 +
# 1792 + 1792 + 4096 per block and channel) and about 10 flops per
# coefficient (576 per block and channel) for the RD terms, 3 channels.
RETILE_FLOPS_PER_PX = 3 * (768 + 2 * 1792 + 4096 + 10 * 576) / 64
# FP32 work of the quant-field refinement per pixel of a refined transform:
# the forward transform (about 16 flops per pixel and channel) and, for each
# of up to 4 candidate field values, the biased quantizer's error and rate
# terms (about 10 flops per coefficient), the inverse transform (about 16 per
# pixel) and the SSIM window sums (about 12 per pixel), 3 channels.
REFINE_FLOPS_PER_PX = 3 * (16 + 4 * (10 + 16 + 12))
VALU_FLOPS_PER_PX = {"select": SELECT_FLOPS_PER_PX, "retile": RETILE_FLOPS_PER_PX,
                     "refine": REFINE_FLOPS_PER_PX}

EXTRA_BYTES_PER_PX = {
    "gaborish": 12.0 + 12.0,  # 3 planes in, 3 planes out
    "epf_proxy_compare": 8.0,  # source + candidate luma (the pyramids are 1/4 and 1/16 of that)
}

CAL_BYTES = 256 << 20  # calibration copy size for the byte counters

# rocprofv3 kernel names of each profiled stage
STAGE_KERNEL = {"xyb": ["k_xyb"], "cmap": ["k_cmap_blocks", "k_cmap_solve"],
                "aq_pre": ["k_aq_pre"], "aq_block": ["k_aq_block"], "select": ["k_select"],
                "quantize": ["k_quantize"], "quantize_xt": ["k_quantize_xt"],
                "fill": ["k_fill_bands"],
                "alpha_prep": ["k_alpha_extract", "k_fringe", "k_fringe_commit", "k_astat_init",
                               "k_alpha_stats", "k_fill_u32", "k_repad", "k_alpha_blocks"],
                "alpha_qf": ["k_alpha_apply"],
                "retile": ["k_piece_costs", "k_retile", "k_adjust_qf"],
                "refine": ["k_zero_u32", "k_refine_list", "k_refine_costs", "k_qf_trellis"]}
# fill_invisible's recurrence: anti-diagonals x + 2 y are its dependent steps,
# and each step's chain is 14 dependent VALU operations (one DPP wave shift,
# 4 fused multiply-adds and 3 adds for the sum, 5 FP64 operations for the
# quotient, one bit select); a wave64 VALU operation takes 4 cycles.
FILL_CHAIN_OPS = 14
CLOCK_GHZ = 2.4  # MI355X peak engine clock

HEADLINE = "smooth_rgba_7680x4320_q75"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_smooth_rgba_7680x4320_q75.txt")
ISSUE_FILE = os.path.join(ROOT, "profiles", "issue_decision_quantize_7680x4320_q75.txt")


def read_traffic(stage):
    """HBM bytes per launch of the stage's kernels from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy)."""
    if not os.path.exists(TRAFFIC_FILE) or stage not in STAGE_KERNEL:
        return None
    got = {}
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "traffic_bytes":
                got[parts[1]] = got.get(parts[1], 0) + int(parts[2])
    names = STAGE_KERNEL[stage]
    if not all(n in got for n in names if n not in ("k_astat_init",)):
        return None
    return sum(got.get(n, 0) for n in names)


def read_valu_busy(kernel):
    """VALU-busy fraction of a kernel from the committed SQ-counter summary."""
    if not os.path.exists(ISSUE_FILE):
        return None
    cur = None
    with open(ISSUE_FILE) as f:
        for line in f:
            if line and not line[0].isspace() and not line.startswith("#"):
                cur = line.split("<")[0].strip()
            elif cur == kernel and line.strip().startswith("VALU busy"):
                return float(line.split()[-1].rstrip("%")) / 100.0
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--quality", type=float, default=75.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile", action="store_true", help="few steps, headline only (for rocprofv3)")
    ap.add_argument("--photo-only", action="store_true",
                    help="run only the 2048x2048 RGB8 photo-like workload (for rocprofv3)")
    ap.add_argument("--batch", type=int, default=512,
                    help="images in the batched-front-end workload (0: skip it)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched workload (for rocprofv3)")
    ap.add_argument("--no-tune", action="store_true", help="skip the quant-field tuning workload")
    ap.add_argument("--tune-only", action="store_true",
                    help="run only the quant-field tuning workload (for rocprofv3)")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(rgb, params, threads):
    """The oracle's whole front end on the host cores (kind "port"), built
    -O3 -march=native for this host (-ffp-contract=off kept: the same bits
    as the checker build).  All cores on the whole image, and one thread on
    a 2048x1080 slab of it, each the median of up to 3 runs after a warm-up."""
    import subprocess
    native = os.path.join(ROOT, "oracle", "_build", "libjxo_native.so")
    try:
        subprocess.check_call(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native"],
                              stdout=subprocess.DEVNULL)
        os.environ["JXO_LIB"] = native
        build = "-O3 -march=native -ffp-contract=off"
    except (OSError, subprocess.CalledProcessError):
        build = "-O2 -march=x86-64-v2 -ffp-contract=off (native build failed)"
    from oracle import oracle as O
    tables = O.tables_for(params)
    h, w, _ = rgb.shape

    def rate(img, n_threads, budget):
        O.front_end(img[:64, :64].copy(), params, tables, n_threads)  # warm
        runs = []
        t_end = time.perf_counter() + budget
        while len(runs) < 3 and (not runs or time.perf_counter() < t_end):
            t0 = time.perf_counter()
            O.front_end(img, params, tables, n_threads)
            runs.append(time.perf_counter() - t0)
        sec = float(np.median(runs))
        return img.shape[0] * img.shape[1] / 1e6 / sec, sec, len(runs)

    all_mp, all_sec, all_n = rate(rgb, threads, 12.0)
    slab = np.ascontiguousarray(rgb[:1080, :2048])
    one_mp, one_sec, one_n = rate(slab, 1, 10.0)
    return {
        "value": round(all_mp, 3),
        "unit": "MP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"whole front end of the same {w}x{h} RGBA image, median of {all_n} runs "
                  f"({all_sec:.2f} s each) with {threads} threads",
        "one_thread": {"value": round(one_mp, 3), "unit": "MP/s",
                       "sample": f"2048x1080 slab of the image, median of {one_n} runs "
                                 f"({one_sec:.2f} s each)"},
        "cpu": cpu_model(),
        "build": build,
    }


def time_kernel(fn, stream, reps, alg_bytes):
    """Mean device time of fn() over reps launches (HIP events on the
    context's stream) and its HBM roofline position."""
    import torch
    fn()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    return {"ms": round(ms, 4), "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "algorithmic_bytes": int(alg_bytes)}


def dct64_forms(fe, xyb, w, h, stream, reps):
    """DCT64X64 of every 64x64 block of the image in its three forms (FP32
    butterflies; split-bf16 x3 and the three-term split x6 on the matrix
    cores), each with its HBM and
    matrix-core roofline position and its error against the FP64 closed form
    on a sample of blocks."""
    import math
    import torch
    nblk = ((w + 63) // 64) * ((h + 63) // 64)
    px = 64 * 64 * nblk
    outs = {m: fe.dct64_image(xyb, w, h, m) for m in (0, 1, 2)}
    k = np.arange(64)[:, None]
    cm = np.cos(np.pi * k * (2 * np.arange(64)[None, :] + 1) / 128)
    cm[1:] *= math.sqrt(2.0)
    planes = xyb.cpu().numpy()
    xbk = 8 * ((w + 63) // 64)
    res = {}
    for m, name in ((0, "dct64_fp32_butterflies"), (1, "dct64_mfma_bf16x3"),
                    (2, "dct64_mfma_bf16x6")):
        r = time_kernel(lambda: fe.dct64_image(xyb, w, h, m, out=outs[m]), stream, reps,
                        24.0 * px)
        # 3 channels, 2 products of 2 * 64^3 flops, 3 or 6 split terms
        flops = 3 * nblk * 2 * (3 if m == 1 else 6) * 2 * 64 ** 3
        if m >= 1:
            r["mfma_tflops"] = round(flops / (r["ms"] * 1e-3) / 1e12, 1)
            r["mfma_frac_bf16_peak"] = round(flops / (r["ms"] * 1e-3) / 2.5e15, 4)
        got = [o.cpu().numpy().reshape(-1, xbk, 64) for o in outs[m]]
        worst = 0.0
        for (tx, ty) in [(0, 0), (3, 5), (7, 2), (12, 9), (31, 31), (20, 27)]:
            if tx * 64 >= w or ty * 64 >= h:
                continue
            for c in range(3):
                x = planes[c, ty * 64:ty * 64 + 64, tx * 64:tx * 64 + 64].astype(np.float64)
                want = (cm @ x @ cm.T / 4096.0).T.reshape(-1)
                g = np.concatenate([got[c][ty * 8 + j, tx * 8:tx * 8 + 8, :].reshape(512)
                                    for j in range(8)])
                worst = max(worst, float(np.abs(g - want).max() / np.abs(want).max()))
        r["max_err_vs_fp64_rel_to_max_coef"] = float("%.3g" % worst)
        res[name] = r
    del outs
    torch.cuda.synchronize()
    return res


def dct128_forms(fe, xyb, w, h, stream, reps):
    """DCT128X128 of every 128x128 block in both forms, as dct64_forms.  The
    planes of a 2048-multiple image already reach whole 128-px blocks."""
    import math
    import torch
    nblk = ((w + 127) // 128) * ((h + 127) // 128)
    px = 128 * 128 * nblk
    outs = {m: fe.dct128_image(xyb, w, h, m) for m in (0, 1)}
    k = np.arange(128)[:, None]
    cm = np.cos(np.pi * k * (2 * np.arange(128)[None, :] + 1) / 256)
    cm[1:] *= math.sqrt(2.0)
    planes = xyb.cpu().numpy()
    xbk = 16 * ((w + 127) // 128)
    res = {}
    for m, name in ((0, "dct128_fp32_butterflies"), (1, "dct128_mfma_bf16x3")):
        r = time_kernel(lambda: fe.dct128_image(xyb, w, h, m, out=outs[m]), stream, reps,
                        24.0 * px)
        flops = 3 * nblk * 2 * 3 * 2 * 128 ** 3
        if m == 1:
            r["mfma_tflops"] = round(flops / (r["ms"] * 1e-3) / 1e12, 1)
            r["mfma_frac_bf16_peak"] = round(flops / (r["ms"] * 1e-3) / 2.5e15, 4)
        got = [o.cpu().numpy().reshape(-1, xbk, 64) for o in outs[m]]
        worst = 0.0
        for (tx, ty) in [(0, 0), (3, 5), (7, 2), (15, 15), (10, 13)]:
            if tx * 128 >= w or ty * 128 >= h:
                continue
            for c in range(3):
                x = planes[c, ty * 128:ty * 128 + 128, tx * 128:tx * 128 + 128].astype(np.float64)
                want = (cm @ x @ cm.T / 16384.0).T.reshape(-1)
                g = np.concatenate([got[c][ty * 16 + j, tx * 16:tx * 16 + 16, :].reshape(1024)
                                    for j in range(16)])
                worst = max(worst, float(np.abs(g - want).max() / np.abs(want).max()))
        r["max_err_vs_fp64_rel_to_max_coef"] = float("%.3g" % worst)
        res[name] = r
    del outs
    torch.cuda.synchronize()
    return res


def host_pipeline_rate(J, torch, fe, d_rgb, params, n_images=6):
    """PCIe-inclusive rate with pinned host buffers: n_images front ends of
    the headline image, input copied from pinned host memory and every output
    (levels, f32 coefficients, DC, field, cells, CfL maps) copied back to
    pinned host memory, the copies of one image overlapped with the next
    image's compute (two device slots, an upload stream and a download
    stream beside the context's).  Never the headline."""
    h, w, ch = d_rgb.shape
    host_in = torch.empty((h, w, ch), dtype=torch.uint8, pin_memory=True)
    host_in.copy_(d_rgb.cpu())
    up, down = torch.cuda.Stream(), torch.cuda.Stream()
    comp = torch.cuda.current_stream()
    slots = []
    for _ in range(2):
        o = fe.outputs(w, h)
        xb, yb = (w + 7) // 8, (h + 7) // 8
        tiles = ((h + 63) // 64, (w + 63) // 64)
        o["qf"] = fe.empty((yb, xb), torch.uint8)
        o["cells"] = fe.empty((yb, xb), torch.uint8)
        o["ytox"] = torch.zeros(tiles, dtype=torch.int8, device="cuda")
        o["ytob"] = torch.zeros(tiles, dtype=torch.int8, device="cuda")
        dev_in = torch.empty_like(d_rgb)
        flat = [o["qf"], o["cells"], o["ytox"], o["ytob"]] + list(o["levels"]) + \
            list(o["coeffs"]) + list(o["dc"])
        host = [torch.empty(t.shape, dtype=t.dtype, pin_memory=True) for t in flat]
        ev = {k: torch.cuda.Event() for k in ("in", "done", "free")}
        ev["free"].record(down)
        slots.append((dev_in, o, flat, host, ev))
    out_bytes = sum(t.numel() * t.element_size() for t in slots[0][2])
    torch.cuda.synchronize()

    def run(n):
        for i in range(n):
            dev_in, o, flat, host, ev = slots[i % 2]
            up.wait_event(ev["free"])
            with torch.cuda.stream(up):
                dev_in.copy_(host_in, non_blocking=True)
            ev["in"].record(up)
            comp.wait_event(ev["in"])
            fe.run(dev_in, w, h, params, outputs=o, qf=o["qf"], cells=o["cells"],
                   ytox=o["ytox"], ytob=o["ytob"])
            ev["done"].record(comp)
            down.wait_event(ev["done"])
            with torch.cuda.stream(down):
                for dst, src in zip(host, flat):
                    dst.copy_(src, non_blocking=True)
            ev["free"].record(down)
        torch.cuda.synchronize()
    run(2)
    t0 = time.perf_counter()
    run(n_images)
    sec = (time.perf_counter() - t0) / n_images
    del slots
    torch.cuda.empty_cache()
    return {"value": round(w * h / 1e6 / sec, 1), "unit": "MP/s", "ms_per_image": round(sec * 1e3, 3),
            "bytes_in_per_image": int(h * w * ch), "bytes_out_per_image": int(out_bytes),
            "note": f"pinned host buffers, {n_images} images, uploads and downloads overlapped "
                    "with the next image's compute; PCIe included"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def batch_workload(J, torch, fe, stream, n, reps):
    """The batch workload: n RGB8 images, sides uniform in [64, 2048] and not
    multiples of 8, quality uniform in [50, 95], photo-like content (seed
    1000 + i, generated on the device).  Autotunes the batch's tile shapes,
    then times graph replays of the whole batch, and the same images as a
    loop of single-image front-end calls."""
    from jixel_amd import synth
    dev = torch.device("cuda", 0)
    sizes = synth.batch_sizes(n, 4)
    field = synth.one_over_f_field(4096, 4, dev)
    images = []
    for i, (w, h, q) in enumerate(sizes):
        img = synth.photo_like_device(w, h, 1000 + i, dev, field)
        images.append((img, J.frame_params(J.EncodeConfig().with_quality(q))))
    del field
    torch.cuda.synchronize()
    px = sum(w * h for w, h, _ in sizes)
    log(f"batch: {n} images, {px / 1e6:.1f} MP generated")
    b = J.Batch(fe, images)
    report = b.autotune(reps=2)
    log("batch: autotuned " + str(b.config()))
    b.launch()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(reps):
        b.launch()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1000.0 / reps
    ms = max(e0.elapsed_time(e1) / reps, wall)
    status = b.status()
    log(f"batch: {ms:.3f} ms per batch")
    # the same images one front-end call each, into the same buffers
    def loop():
        for i, (img, p) in enumerate(images):
            o = b.outputs(i)
            fe.run(img, img.shape[1], img.shape[0], p, outputs=o, qf=o["qf"], cells=o["cells"],
                   ytox=o["ytox"], ytob=o["ytob"])
    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    loop()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_loop = max(e0.elapsed_time(e1), (time.perf_counter() - t0) * 1000.0)
    log(f"batch: single-image loop {ms_loop:.3f} ms")
    out = {
        "workload": f"batch_{n}_photo_like_rgb8_sides_64_2048_q50_95",
        "images": n,
        "megapixels": round(px / 1e6, 2),
        "ms_per_batch": round(ms, 3),
        "mp_per_s": round(px / 1e6 / (ms * 1e-3), 1),
        "single_image_loop_ms": round(ms_loop, 3),
        "single_image_loop_mp_per_s": round(px / 1e6 / (ms_loop * 1e-3), 1),
        "tile_config": b.config(),
        "autotune_ms": {line.split(" ")[0] + " " + line.split(" ")[1] + " " + line.split(" ")[2]:
                        float(line.split(" ")[3]) for line in report.strip().splitlines()},
        "launches_per_batch": 8,
        "graph": True,
        "status": list(status),
        "data": "synthetic, generated on the device (photo-like recipe, seeds 1000 + i)",
    }
    b.close()
    del images
    torch.cuda.empty_cache()
    return out


def tune_workload(J, torch, fe, stream, n=16, side=1024, steps=20, lr=0.05, lam=0.01):
    """Quant-field tuning: n photo-like side x side RGB8 images at q75 through
    the front end, then `steps` Adam steps of the gradient tuner per image
    (lambda, lr as given) and the rounded integer field.  Reports the device
    time of the tuning (forward + adjoint + Adam) and the proxy loss before
    and after."""
    from jixel_amd import synth
    dev = torch.device("cuda", 0)
    field = synth.one_over_f_field(4096, 5, dev)
    p = J.frame_params(J.EncodeConfig().with_quality(75.0))
    jobs = []
    for i in range(n):
        rgb = synth.photo_like_device(side, side, 5000 + i, dev, field)
        o = fe.run(rgb, side, side, p)
        xyb = fe.xyb(rgb, side, side)
        jobs.append((xyb, o["cells"], o["qf"]))
    del field
    torch.cuda.synchronize()
    log(f"tune: {n} images through the front end")
    tuners = [J.QfTuner(fe, xyb, side, side, cells, p, lam=lam, seed=7, max_steps=steps)
              for xyb, cells, _ in jobs]
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    traces = []
    for t, (_, _, qf) in zip(tuners, jobs):
        _, _, trace = t.run(qf, steps=steps, lr=lr)
        traces.append(trace)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = max(e0.elapsed_time(e1), (time.perf_counter() - t0) * 1000.0)
    log(f"tune: {ms:.2f} ms")
    for t in tuners:
        t.close()
    before = sum(tr[0] for tr in traces)
    after = sum(tr[-1] for tr in traces)
    px = n * side * side
    return {
        "workload": f"qf_tune_{n}x{side}x{side}_photo_like_q75",
        "images": n, "steps": steps, "lr": lr, "lambda": lam,
        "ms_total": round(ms, 3),
        "ms_per_adam_step_per_mp": round(ms / steps / (px / 1e6), 4),
        "mp_per_s_per_step": round(px / 1e6 * steps / (ms * 1e-3), 1),
        "loss_before": round(before, 2),
        "loss_after": round(after, 2),
        "note": "proxy loss (channel-weighted squared error + lambda * log-rate); the "
                "SSIMULACRA2-at-equal-bytes target needs the CPU back end and is not measured",
    }


def front_end_workload(J, torch, fe, stream, d_rgb, p, steps, warmup):
    """Times `steps` front-end calls on a device-resident image (HIP events on
    the context's stream bracketing the whole timed region, and the host
    clock), with the library's per-kernel event pairs recorded inside that
    region.  Returns (ms per step, {stage: mean kernel ms})."""
    h, w, _ = d_rgb.shape
    dev = d_rgb.device
    outs = fe.outputs(w, h)
    xb, yb = (w + 7) // 8, (h + 7) // 8
    qf = fe.empty((yb, xb), torch.uint8)
    cells = fe.empty((yb, xb), torch.uint8)
    tiles = ((h + 63) // 64, (w + 63) // 64)
    ytox = torch.zeros(tiles, dtype=torch.int8, device=dev)
    ytob = torch.zeros(tiles, dtype=torch.int8, device=dev)

    def step():
        fe.run(d_rgb, w, h, p, outputs=outs, qf=qf, cells=cells, ytox=ytox, ytob=ytob)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    fe.profile(True)
    t0 = time.perf_counter()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = fe.profile_read()
    fe.profile(False)
    ms_gpu = e0.elapsed_time(e1) / steps
    ms_step = max(ms_gpu, wall * 1000.0 / steps)
    del outs, qf, cells, ytox, ytob
    k = {n: v[0] for n, v in prof.items() if v[1] > 0}
    if "alpha" in k and "fill" in k:  # the alpha stage without its fill_invisible kernel
        k["alpha_prep"] = k.pop("alpha") - k["fill"]
    return ms_step, k


def roofline_of(kernel_ms, px_shape, traffic_stage_ok=True):
    """Roofline of the dominant kernel.  The decision kernels (strategy
    search, re-tiling, quant-field refinement) keep their tiles in LDS and
    are bound by VALU issue: their bound is FP32 vector throughput, with the
    HBM rate kept as a secondary figure.  The others are HBM-bound."""
    dominant = max(kernel_ms, key=kernel_ms.get)
    px = px_shape[0] * px_shape[1]
    sec = kernel_ms[dominant] * 1e-3
    alg_bytes = KERNEL_BYTES_PER_PX[dominant] * px
    gbs = alg_bytes / sec / 1e9
    hbm = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 4)}
    r = {"kernel": dominant, "bound": "hbm"}
    r.update(hbm)
    r["traffic"] = read_traffic(dominant) if traffic_stage_ok else None
    r["algorithmic_bytes_per_launch"] = int(alg_bytes)
    r["kernel_ms"] = round(kernel_ms[dominant], 4)
    if dominant == "fill":
        # the recurrence's dependent steps bound it, not HBM: the anti-diagonals
        # times the per-step chain at 4 cycles per dependent operation
        w, h = px_shape
        steps = w + 2 * (h - 1)
        ideal = steps * FILL_CHAIN_OPS * 4 / (CLOCK_GHZ * 1e9) * 1e3
        r["limiter"] = ("the recurrence's W + 2 H dependent steps, not HBM (3 x 68 bands of 64 "
                        "rows, one workgroup each)")
        r["latency_bound"] = {"anti_diagonals": steps, "dependent_ops_per_step": FILL_CHAIN_OPS,
                              "cycles_per_op": 4, "clock_ghz": CLOCK_GHZ,
                              "ideal_ms": round(ideal, 4),
                              "frac": round(ideal / kernel_ms[dominant], 4)}
    if dominant in VALU_FLOPS_PER_PX:
        fpp = VALU_FLOPS_PER_PX[dominant]
        tflops = fpp * px / sec / 1e12
        r.update({"bound": "valu", "achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS,
                  "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
                  "flops_per_px": round(fpp, 1), "hbm": hbm})
        r["limiter"] = "VALU issue, not HBM: the candidate transforms and RD sums run in LDS"
        # the stage's main kernel in the committed issue-counter summary
        main = {"select": "k_select", "retile": "k_piece_costs", "refine": "k_refine_costs"}
        r["valu_busy"] = read_valu_busy(main[dominant]) if traffic_stage_ok else None
        r["valu_busy_kernel"] = main[dominant]
    return r


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; every box here has one "
                                   "MI355X and the front end shards by image / DC group "
                                   "without collectives (see DESIGN.md)"}))
        sys.exit(2)
    import torch

    import jixel_amd as J
    from jixel_amd import synth

    if args.profile:
        args.steps, args.warmup, args.no_cpu = min(args.steps, 5), min(args.warmup, 2), True
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    fe = J.FrontEnd(0, stream=stream)
    if args.tune_only:
        print(json.dumps({"qf_tune": tune_workload(J, torch, fe, stream)}), flush=True)
        fe.close()
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_workload(J, torch, fe, stream, max(1, args.batch),
                                                  max(1, min(args.steps, 3)))}), flush=True)
        fe.close()
        return
    if args.profile:
        # Calibration for the HBM byte counters: one plain copy kernel that
        # reads and writes exactly CAL_BYTES each (a 256 MiB f32 negation).
        cal = torch.ones(CAL_BYTES // 4, dtype=torch.float32, device=dev)
        cal_out = torch.empty_like(cal)
        torch.cuda.synchronize()
        torch.neg(cal, out=cal_out)
        torch.cuda.synchronize()
        del cal, cal_out
    cfg = J.EncodeConfig().with_quality(args.quality)
    p = J.frame_params(cfg)
    # the 2048x2048 RGB8 photo-like workload
    photo = synth.photo_like(2048, 2048, 2)
    d_photo = torch.from_numpy(photo).to(dev)
    photo_ms, photo_k = front_end_workload(J, torch, fe, stream, d_photo, p,
                                           args.steps if args.photo_only else max(5, args.steps // 2),
                                           args.warmup)
    photo_res = {
        "workload": "photo_like_2048x2048_rgb8_q75",
        "mp_per_s": round(2048 * 2048 / 1e6 / (photo_ms * 1e-3), 2),
        "ms_per_step": round(photo_ms, 4),
        "kernel_ms": {k: round(v, 4) for k, v in photo_k.items()},
        "roofline": roofline_of(photo_k, (2048, 2048), traffic_stage_ok=False),
    }
    # the same photo at q90 (d = 1.0): the decision adds the SSIM rerank and
    # drops the quant-field refinement
    p90 = J.frame_params(J.EncodeConfig().with_quality(90.0))
    q90_ms, q90_k = front_end_workload(J, torch, fe, stream, d_photo, p90,
                                       max(3, args.steps // 4), args.warmup)
    photo_res["q90"] = {
        "workload": "photo_like_2048x2048_rgb8_q90",
        "distance": round(p90.distance, 4),
        "mp_per_s": round(2048 * 2048 / 1e6 / (q90_ms * 1e-3), 2),
        "ms_per_step": round(q90_ms, 4),
        "kernel_ms": {k: round(v, 4) for k, v in q90_k.items()},
    }
    del d_photo
    if args.photo_only:
        print(json.dumps({"photo": photo_res}), flush=True)
        fe.close()
        return
    # the headline: 7680x4320 RGBA q75
    w, h = 7680, 4320
    log(f"generating the {w}x{h} RGBA image on the device")
    d_rgb = synth.smooth_rgba_device(w, h, 3, dev)
    torch.cuda.synchronize()
    ms_step, kernel_ms = front_end_workload(J, torch, fe, stream, d_rgb, p, args.steps, args.warmup)
    log(f"front end {w}x{h}: {ms_step:.3f} ms per step")
    px = w * h

    # Kernels beside the measured chain (not part of the Fast-tier step),
    # each timed alone on the same image with its algorithmic bytes.
    extra = {}
    reps = max(3, min(args.steps, 10))
    xyb = fe.xyb(d_rgb, w, h, fused=bool(p.fused))
    gab_out = torch.empty_like(xyb)

    def gab():
        fe.gaborish_inverse(xyb, w, h, fused=bool(p.fused), out=gab_out)
    extra["gaborish"] = time_kernel(gab, stream, reps, EXTRA_BYTES_PER_PX["gaborish"] * px)
    # EPF proxy: source = XYB luma, candidate = the gaborish-sharpened luma
    proxy = J.LumaProxy(fe, xyb[1, :h, :w])
    cand = gab_out[1, :h, :w]
    extra["epf_proxy_compare"] = time_kernel(lambda: proxy.compare(cand), stream, reps,
                                             EXTRA_BYTES_PER_PX["epf_proxy_compare"] * px)
    extra["epf_proxy_compare"]["note"] = "host-synchronous call (score read back)"
    proxy.close()
    del gab_out
    if not args.profile:
        extra.update(dct64_forms(fe, xyb, w, h, stream, reps))
        extra.update(dct128_forms(fe, xyb, w, h, stream, reps))
    del xyb
    torch.cuda.empty_cache()

    result = {
        "metric": "VarDCT front-end MP/s at q75 (1 GPU vs CPU); ssimulacra2 at equal bytes",
        "value": round(px / 1e6 / (ms_step * 1e-3), 2),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (1/f noise smoothed over 33-129 px, 20 soft-edged alpha ellipses; "
                "generated on the device, seed 3)",
        "config": {
            "workload": HEADLINE,
            "width": w, "height": h, "channels": 4, "quality": args.quality,
            "distance": round(p.distance, 4), "speed": "fast",
            "decisions": "square merges DCT8/DCT16X16/DCT32X32, layout_dp::retile_coefficient "
                         "(aligned + (4,4)-shifted grids, unaligned pairs and squares), "
                         "adjust_quant_field, refine_quant_field with qf_trellis "
                         "(2 steps at d >= 2); the SSIM rerank runs at d <= 1.709: "
                         "the FAST tier's whole decision",
            "alpha": "alpha_fringe::clear, fill_invisible, alpha block factors",
            "cfl": "per-tile regression maps (fill_cmap, Fast path)",
            "tables": "DequantMatrices::new_fast tier 3 (SS2 scales, flat-B DCT8 row)",
            "fmla": "fused" if p.fused else "unfused",
            "input": "u8 RGBA resident in HBM; outputs (levels, f32 coefficients, DC, field, "
                     "strategy, CfL maps) left in HBM",
        },
        "roofline": roofline_of(kernel_ms, (w, h)),
        "kernel_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
        "other_kernels": extra,
        "photo_2048": photo_res,
    }
    if not args.profile:
        result["host_pinned_pipeline"] = host_pipeline_rate(J, torch, fe, d_rgb, p)
        if args.batch > 0:
            result["batch"] = batch_workload(J, torch, fe, stream, args.batch, 3)
        if not args.no_tune:
            result["qf_tune"] = tune_workload(J, torch, fe, stream)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        result["cpu_baseline"] = cpu_baseline(d_rgb.cpu().numpy(), p, threads)
    else:
        result["cpu_baseline"] = None
    fe.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
