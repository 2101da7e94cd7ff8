"""Benchmark of the MI355X render path on one GPU.

Workload (the headline configuration, C2): a 512^3 fp32 volume (4-octave
Perlin fBm + 32 Gaussian blobs, zeroed below its 40th percentile) stored as
2x2x2 bricks of 256^3 with a one-voxel ghost layer, rendered to a 1024x1024
RGBA8 frame with a 256-entry transfer function, 0.5-voxel steps, sort-last
compositing of the per-brick partials (early ray termination at alpha >= 0.99
inside each brick, "over" far to near).  Volume, LUT and frame are resident in
HBM before timing; a step is one full frame.

value = Gsamples/s = S / device time per frame, S = samples classified and
composited (after macrocell skipping, up to early termination).
roofline: HBM bound; algorithmic bytes per frame = S * 8 * sizeof(voxel) (the
trilinear footprint with no reuse) + 4 B per output pixel (+ 32 B per tile
pixel when tiles round-trip HBM), divided by the average device time of a
frame (HIP events on the library's stream), against 8 TB/s.  traffic: the HBM
bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE for the same
command, scaled by the counters' reading for a copy of known size
(profiles/traffic.json).
cpu_baseline: the oracle (the CPU restatement of the same workload, same
samples and arithmetic, one pass per ray) on the host cores (the CPUs of the
process's affinity mask, capped by OMP_NUM_THREADS), whole frames.
reference_cpu: the reference ray.c renderer compiled from its sources
(oracle/_ref) with its own OpenMP loop on the same thread count and on one
core, against this library's compat mode on the GPU over the same volume,
camera and LUT (bit-identical images).

--config 4 (learned surrogate of the C2 volume), --config 5 (steering
service) and --gradient (the central-difference gradient field of the C2
volume) print their own JSON lines; they are measurements of those parts,
not the headline.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--skip", type=int, default=-1, help="macrocell skipping (default per config)")
    ap.add_argument("--composite", choices=["fused", "sort_last"], default=None,
                    help="brick compositing (default per config)")
    ap.add_argument("--no-cpu", action="store_true", help="leave out the CPU legs")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--train-seconds", type=float, default=60.0, help="--config 4/5: training time budget")
    ap.add_argument("--sg-volume", type=int, default=2, choices=[2, 3],
                    help="--config 4: train on the C2 (512^3 fp32) or the C3 (1024^3 fp16) volume")
    ap.add_argument("--serve-surrogate", action="store_true",
                    help="--config 5: also serve the 1024^3 volume from a surrogate trained for --train-seconds")
    ap.add_argument("--requests", type=int, default=400, help="--config 5: requests per arrival rate")
    ap.add_argument("--gradient", action="store_true",
                    help="measure the gradient-field pass over the C2 volume instead of rendering")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (for scaling PMC byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build renders all bricks on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    # every OpenMP CPU leg (the oracle, the reference) runs on the same thread count
    os.environ["OMP_NUM_THREADS"] = str(cpu_threads())
    import numpy as np
    import ivr_amd as ivr
    from ivr_amd import configs

    if args.gradient:
        bench_gradient(args)
        return
    if args.config == 4:
        bench_surrogate(args)
        return
    if args.config == 5:
        bench_serving(args)
        return
    c = args.config
    cfg = configs.WORKLOADS[c]
    W, H = cfg["image"]
    skip = cfg["skip"] if args.skip < 0 else bool(args.skip)
    names = {"fused": ivr.FUSED, "sort_last": ivr.SORT_LAST}
    comp = cfg["composite"] if args.composite is None else names[args.composite]
    comp_name = {v: k for k, v in names.items()}[comp]
    ctx = ivr.Context(0)
    log("generating", cfg["name"])
    vol = ctx.volume(*cfg["dims"], cfg["dtype"], brick=cfg["brick"])
    t0 = time.time()
    vol.generate(configs.gen_params(c))
    log("generated in %.2fs" % (time.time() - t0))
    lut = configs.tf_lut(c)
    tf = ctx.tf(lut)
    st = configs.state(c)
    desc = ivr.config_desc(st, skip=skip, composite=comp)
    fr = ctx.frame(W, H)

    def timed(d, steps, warmup):
        for _ in range(warmup):
            ctx.render(vol, tf, d, fr)
        ctx.sync()
        stats = fr.stats()
        ctx.sync()
        t_wall = time.perf_counter()
        ctx.record(0)
        for _ in range(steps):
            ctx.render(vol, tf, d, fr)
        ctx.record(1)
        ctx.sync()
        t_wall = time.perf_counter() - t_wall
        return ctx.elapsed(0, 1) / steps, t_wall * 1e3 / steps, stats

    dev_ms, wall_ms, stats = timed(desc, args.steps, args.warmup)
    S = stats["visited"]  # classified samples (equal to stats["samples"] without skipping)
    gsps = S / (dev_ms * 1e-3) / 1e9
    log("%s: %d samples classified per frame (%d incl. skipped); %.3f ms/frame device, %.3f wall, "
        "%.1f Gsamples/s" % (comp_name, S, stats["samples"], dev_ms, wall_ms, gsps))
    img = fr.download()
    # the dominant kernel's own launch duration, from the HIP events the frame
    # records on the context's stream around it (and around the combine pass)
    k_ms, c_ms = kernel_ms(ctx, lambda: ctx.render(vol, tf, desc, fr), fr, args.steps)
    vox_bytes = 2 if cfg["dtype"] == ivr.IVR_F16 else 4
    vt = "float" if vox_bytes == 4 else "_Float16"
    if comp == ivr.SORT_LAST:
        # one workgroup per (pixel block, run): 8 voxels per classified sample
        # read, one RGBA32F partial plane per run written
        runs = sum(vol.info()["nbricks"]) - 2
        # more than 4 run slots: per-ray run counts, only the existing partials written
        counts = runs > 4
        bytes_alg = S * 8 * vox_bytes + stats["partials"] * 16 + (4 * W * H if counts else 0)
        kernel = "k_runs<%s, %s, %s>" % (vt, str(skip).lower(), str(counts).lower())
    else:
        bytes_alg = S * 8 * vox_bytes + 4 * W * H
        kernel = "k_render<1, %s, %s, 0>" % (vt, str(skip).lower())
    achieved = bytes_alg / (k_ms * 1e-3)
    traffic = measured_traffic(cfg["name"], comp_name, skip)

    other = {}
    alt = ivr.FUSED if comp != ivr.FUSED else ivr.SORT_LAST
    a_ms, _, a_stats = timed(ivr.config_desc(st, skip=skip, composite=alt), args.steps, 2)
    other[{v: k for k, v in names.items()}[alt]] = {
        "ms_per_frame": round(a_ms, 4), "samples_per_frame": int(a_stats["visited"]),
        "gsamples_s": round(a_stats["visited"] / a_ms / 1e6, 2)}
    ctx.render(vol, tf, desc, fr)
    # the host-buffer boundary: each frame rendered and its RGBA8 image copied
    # over PCIe into pageable host memory before the next one is queued, as the
    # frame protocol's reply needs it, into one reused buffer as a host's colour
    # buffer is (wall clock; not the reported value)
    host_img = fr.download()
    t_host = time.perf_counter()
    for _ in range(args.steps):
        ctx.render(vol, tf, desc, fr)
        fr.download(out=host_img)
    host_ms = (time.perf_counter() - t_host) * 1e3 / args.steps

    cpu = ref = None
    if not args.no_cpu:
        cpu = cpu_baseline(vol, st, lut, img, W, H, comp, cfg["brick"], args.cpu_seconds)
        ref = reference_cpu(ctx, vol, st, lut, W, H)

    out = {
        "metric": "Gsamples/s, ms/frame 512³→1024²; surrogate fps @PSNR≥40dB; p99 latency",
        "value": round(gsps, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dev_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if cfg["dtype"] == ivr.IVR_F32 else "f16 storage, f32 compute",
        "data": "synthetic (seeded Perlin fBm + Gaussian blobs, generated on the GPU)",
        "config": {
            "workload": cfg["name"],
            "image": "%dx%d" % (W, H),
            "composite": comp_name,
            "macrocell_skip": bool(skip),
            "samples_per_frame": int(S),
            "samples_incl_skipped": int(stats["samples"]),
            "step_voxels": configs.STEP_VOX,
            "ert": configs.ERT,
            "frames_per_s": round(1000.0 / dev_ms, 2),
            "wall_ms_per_frame": round(wall_ms, 4),
            "with_rgba8_download": {"ms_per_frame": round(host_ms, 4),
                                    "gsamples_s": round(S / host_ms / 1e6, 3)},
            "other_composite": other,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "kernel_ms": round(k_ms, 4),
            "combine_ms": round(c_ms, 4),
            "achieved": round(achieved / 1e9, 1),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "reference_cpu": ref,
    }
    if cpu and cpu["value"] > 0:
        out["config"]["gpu_over_cpu"] = round(gsps / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    fr.close()
    tf.close()
    vol.close()
    ctx.close()


def kernel_ms(ctx, render, fr, steps):
    """mean duration of the frame's render kernel and of its combine pass
    (zero when there is none) over `steps` frames, each read back from the
    frame's HIP events before the next frame is queued"""
    r, c = [], []
    for _ in range(steps):
        render()
        s = fr.stats()
        r.append(s["render_ms"])
        c.append(s["composite_ms"])
    return sum(r) / len(r), sum(c) / len(c)


def _frame_ms(ctx, fn, steps):
    fn()
    ctx.sync()
    ctx.record(0)
    for _ in range(steps):
        fn()
    ctx.record(1)
    ctx.sync()
    return ctx.elapsed(0, 1) / steps


def train_surrogate(ctx, vol, seconds, batch=1 << 24, chunk=5):
    """Adam steps of `batch` samples in chunks until `seconds` of training"""
    sg = ctx.surrogate(vol)
    losses, steps, t_train = [], 0, 0.0
    while t_train < seconds:
        ctx.sync()
        t = time.perf_counter()
        losses.append(sg.train(vol, chunk, batch))
        ctx.sync()
        t_train += time.perf_counter() - t
        steps += chunk
        log("surrogate: %d steps, %.1fs, loss %.3e" % (steps, t_train, losses[-1]))
    return sg, losses, steps, t_train


def bench_surrogate(args):
    """C4: per-brick hash grids + BF16 MFMA MLP trained on the C2 volume
    (or, with --sg-volume 3, the 1024^3 fp16 C3 volume) with 2^24 samples
    per Adam step for a time budget, then 1920x1080 frames from the
    surrogate against the plain render: PSNR/SSIM, frames/s, and the
    per-brick fallback below 40 dB."""
    import numpy as np
    import ivr_amd as ivr
    from ivr_amd import configs
    c = args.sg_volume
    cfg = configs.WORKLOADS[c]
    ctx = ivr.Context(0)
    vol = ctx.volume(*cfg["dims"], cfg["dtype"], brick=cfg["brick"])
    vol.generate(configs.gen_params(c))
    W, H = 1920, 1080
    lut = configs.tf_lut(c)
    tf = ctx.tf(lut)
    desc = ivr.config_desc(ivr.orbit_state(3004, W, H))  # camera seed of C4
    fr = ctx.frame(W, H)
    plain_ms = _frame_ms(ctx, lambda: ctx.render(vol, tf, desc, fr), args.steps)
    plain = fr.download()
    batch = 1 << 24
    sg, losses, steps, t_train = train_surrogate(ctx, vol, args.train_seconds, batch)
    step_ms = t_train * 1e3 / steps
    sg.set_fallback(np.zeros(sg.nbricks, np.uint8))
    ns_ms = _frame_ms(ctx, lambda: sg.render(vol, tf, desc, fr), args.steps)
    p_ns = ivr.psnr(fr.download(), plain)
    desc = ivr.config_desc(desc.state, skip=True)  # the network sees only samples of non-transparent cells
    sg_ms = _frame_ms(ctx, lambda: sg.render(vol, tf, desc, fr), args.steps)
    sg_stats = fr.stats()
    img = fr.download()
    p_sg, s_sg = ivr.psnr(img, plain), ivr.ssim(img, plain)
    bpsnr, nfall = sg.evaluate(vol, tf, desc, threshold_db=40.0)
    fb_ms = _frame_ms(ctx, lambda: sg.render(vol, tf, desc, fr), args.steps)
    img_fb = fr.download()
    p_fb, s_fb = ivr.psnr(img_fb, plain), ivr.ssim(img_fb, plain)
    log("surrogate: train %.1f ms/step (%d steps), render %.3f ms (plain %.3f); PSNR %.2f dB SSIM %.4f; "
        "%d/%d bricks below 40 dB -> with fallback PSNR %.2f dB, %.3f ms"
        % (step_ms, steps, sg_ms, plain_ms, p_sg, s_sg, nfall, sg.nbricks, p_fb, fb_ms))
    flops_step = batch * 3 * 2 * (32 * 64 + 64 * 64 + 64)
    workload = "C4 surrogate of %s: %d bricks x 16 levels x 2^19 x 2 features, MLP 32-64-64-1" \
        % ("C2" if c == 2 else "C3 (1024^3 fp16)", sg.nbricks)
    # roofline of the network kernel: per network sample 16 levels x 8 corners x 2 half-precision
    # features gathered (512 B) and 2 x (32x64 + 64x64 + 64) = 12,416 flops, over the device time of
    # the frame's macrocell pass + k_render_sg (HIP events)
    k_ms, _ = kernel_ms(ctx, lambda: sg.render(vol, tf, desc, fr), fr, args.steps)
    S = int(fr.stats()["visited"])
    g_bytes = S * 16 * 8 * 4
    flops = S * 2 * (32 * 64 + 64 * 64 + 64)
    roofline = {"bound": "hbm", "kernel": "k_render_sg<true>", "kernel_ms": round(k_ms, 4),
                "achieved": round(g_bytes / (k_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(g_bytes / (k_ms * 1e-3) / HBM_PEAK, 4),
                "traffic": measured_traffic(workload, "surrogate", True),
                "mfma_tflops": round(flops / (k_ms * 1e-3) / 1e12, 2), "mfma_peak_tflops": 2500.0,
                "mfma_frac": round(flops / (k_ms * 1e-3) / 2.5e15, 5)}
    cpu = None if args.no_cpu else cpu_surrogate_baseline(sg, args.cpu_seconds)
    out = {
        "metric": "surrogate frames/s at 1920x1080 with per-brick fallback below 40 dB",
        "value": round(1000.0 / fb_ms, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16 MLP on MFMA, fp32 tables and accumulation",
        "data": "synthetic %s volume (%s), surrogate trained on the GPU"
                % ("C2" if c == 2 else "C3", "512^3 fp32" if c == 2 else "1024^3 fp16"),
        "config": {
            "workload": workload,
            "image": "%dx%d" % (W, H),
            "train_steps": steps, "train_samples_per_step": batch, "train_ms_per_step": round(step_ms, 2),
            "train_mlp_tflops": round(flops_step / (step_ms * 1e-3) / 1e12, 2),
            "loss_first": losses[0], "loss_last": losses[-1],
            "surrogate_ms": round(sg_ms, 3), "surrogate_fps": round(1000.0 / sg_ms, 2),
            "surrogate_no_skip_ms": round(ns_ms, 3), "surrogate_no_skip_psnr_db": round(p_ns, 2),
            "network_samples_per_frame": int(sg_stats["visited"]),
            "network_gsamples_s": round(sg_stats["visited"] / sg_ms / 1e6, 3),
            "psnr_db": round(p_sg, 2), "ssim": round(s_sg, 4),
            "brick_psnr_db": [round(float(x), 2) for x in bpsnr],
            "fallback_bricks": int(nfall),
            "fallback_ms": round(fb_ms, 3), "fallback_psnr_db": round(p_fb, 2), "fallback_ssim": round(s_fb, 4),
            "plain_ms": round(plain_ms, 3), "plain_fps": round(1000.0 / plain_ms, 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    sg.close()
    fr.close()
    tf.close()
    vol.close()
    ctx.close()


def cpu_surrogate_baseline(sg, target_s):
    """The oracle's forward-only surrogate evaluation (hash encoding + MLP, fp64
    with the bf16 rounding points) of uniform points in brick 0 on the host
    cores, sized to about target_s seconds: the per-sample network work of a
    surrogate frame, without the march."""
    import numpy as np
    import oracle
    mlp, table = sg.get(brick=0, table=True)
    threads = cpu_threads()
    rng = np.random.default_rng(4004)
    n = 1 << 12
    t = time.perf_counter()
    oracle.sg_eval(mlp, table, sg.T, rng.uniform(0.0, 1.0, (n, 3)))
    dt = time.perf_counter() - t
    n = int(min(1 << 24, max(n, n * target_s / max(dt, 1e-4))))
    xyz = rng.uniform(0.0, 1.0, (n, 3))
    t = time.perf_counter()
    oracle.sg_eval(mlp, table, sg.T, xyz)
    dt = time.perf_counter() - t
    log("cpu surrogate baseline: %d points in %.2fs on %d threads" % (n, dt, threads))
    return {"value": round(n / dt / 1e6, 3), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "host": host_info(),
            "sample": "%d uniform points of brick 0 through the oracle's encoding + MLP (fp64, bf16 "
                      "rounding), one timed run" % n}


def bench_serving(args):
    """C5: three resident volumes (256^3 fp32, C2 512^3 fp32, C3 1024^3 fp16)
    behind the continuous batcher; Poisson arrivals at 50..400 req/s and an
    all-at-once burst for capacity; each request a seeded orbit camera, a
    jittered TF and a resolution from {256^2, 512^2, 1024^2, 1920x1080}."""
    import ivr_amd as ivr
    from ivr_amd import configs, serving
    import numpy as np
    ctx = ivr.Context(0)
    vols = []
    for (n, dt, c, brick) in ((256, ivr.IVR_F32, 2, 256), (512, ivr.IVR_F32, 2, 256), (1024, ivr.IVR_F16, 3, 256)):
        v = ctx.volume(n, n, n, dt, brick=brick)
        v.generate(configs.gen_params(c))
        vols.append(v)

    def sweep(bt, tag):
        rows = []
        for k, rate in enumerate((50.0, 100.0, 200.0, 400.0, 1e9)):
            reqs = serving.make_requests(args.requests, len(vols), seed=4100 + k, tf_seed=2005)
            res = serving.run_poisson(bt, reqs, rate, seed=4000 + k)
            row = {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in res.items() if kk != "records"}
            row["rate"] = "burst" if rate >= 1e9 else rate
            # latency of the requests of the 1024^3 resident (volume 2) apart
            lat3 = [(r["t_done"] - r["t_arrival"]) * 1e3 for r in res["records"].values()
                    if r["status"] == 0 and reqs[r["id"] - 1]["volume"] == 2]
            if lat3:
                row["vol1024_p50_ms"] = round(float(np.percentile(lat3, 50)), 3)
                row["vol1024_p99_ms"] = round(float(np.percentile(lat3, 99)), 3)
            log("serving (%s): %s" % (tag, row))
            rows.append(row)
        return rows

    bt = ctx.batcher(64 << 30, 16, 1920, 1080)
    for v in vols:
        bt.add_volume(v)
    rows = sweep(bt, "plain")
    st = bt.stats()
    bt.close()
    sg_part = None
    if args.serve_surrogate:
        # the 1024^3 resident served from its surrogate (fallback bricks from the plain volume)
        sg, losses, steps, t_train = train_surrogate(ctx, vols[2], args.train_seconds)
        lut = configs.tf_lut(3)
        tf = ctx.tf(lut)
        desc = ivr.config_desc(ivr.orbit_state(3004, 1920, 1080), skip=True)
        bpsnr, nfall = sg.evaluate(vols[2], tf, desc, threshold_db=40.0)
        tf.close()
        bt = ctx.batcher(64 << 30, 16, 1920, 1080)
        bt.add_volume(vols[0])
        bt.add_volume(vols[1])
        bt.add_surrogate(sg, vols[2])
        srows = sweep(bt, "surrogate")
        bt.close()
        sg_part = {"train_steps": steps, "train_s": round(t_train, 1), "loss_last": losses[-1],
                   "fallback_bricks": int(nfall), "bricks": sg.nbricks,
                   "brick_psnr_db_min": round(float(bpsnr.min()), 2), "rates": srows,
                   "capacity_frames_per_s": srows[-1]["frames_per_s"]}
        sg.close()
    out = {
        "metric": "steering service p99 latency (arrival -> RGBA8 on host) and frames/s",
        "value": rows[-2]["p99_ms"],
        "unit": "ms (p99 at 400 req/s)",
        "n_gpus": 1,
        "higher_is_better": False,
        "data": "synthetic volumes generated on the GPU; seeded requests",
        "config": {"workload": "C5: 256^3 fp32, 512^3 fp32, 1024^3 fp16 resident; 16 slots; 64 GiB budget",
                   "requests_per_rate": args.requests, "rates": rows,
                   "capacity_frames_per_s": rows[-1]["frames_per_s"],
                   "batcher": {k: v for k, v in st.items()} if isinstance(st, dict) else None,
                   "surrogate_1024": sg_part},
    }
    print(json.dumps(out), flush=True)
    for v in vols:
        v.close()
    ctx.close()


def bench_gradient(args):
    """--gradient: ivr_volume_gradient over the C2 volume (512^3 fp32, 2x2x2
    bricks) into device memory.  HBM bound: 4 B read + 16 B written per voxel.
    Checked bit for bit against the oracle over the whole volume; the oracle's
    gradient on the host cores is the CPU baseline."""
    import ctypes as C
    import numpy as np
    import ivr_amd as ivr
    from ivr_amd import configs
    import oracle
    cfg = configs.WORKLOADS[2]
    nx, ny, nz = cfg["dims"]
    nvox = nx * ny * nz
    ctx = ivr.Context(0)
    vol = ctx.volume(nx, ny, nz, cfg["dtype"], brick=cfg["brick"])
    vol.generate(configs.gen_params(2))
    hip = C.CDLL("libamdhip64.so")
    dst = C.c_void_p()
    if hip.hipMalloc(C.byref(dst), C.c_size_t(nvox * 16)) != 0:
        raise RuntimeError("hipMalloc of the gradient field failed")
    fn = lambda: vol.gradient(device_ptr=dst.value)  # noqa: E731
    for _ in range(args.warmup):
        fn()
    ctx.sync()
    ms = _frame_ms(ctx, fn, args.steps)
    got = np.empty((nz, ny, nx, 4), np.float32)
    if hip.hipMemcpy(C.c_void_p(got.ctypes.data), dst, C.c_size_t(nvox * 16), 2) != 0:  # D2H
        raise RuntimeError("hipMemcpy of the gradient field failed")
    hip.hipFree(dst)
    host = vol.download()
    t = time.perf_counter()
    ref = oracle.gradient(host)
    cpu_s = time.perf_counter() - t
    same = bool(np.array_equal(got.view(np.uint32), ref.view(np.uint32)))
    log("gradient: %.4f ms per pass, %.1f Gvoxels/s; bit-identical to the oracle: %s; oracle %.2fs"
        % (ms, nvox / ms / 1e6, same, cpu_s))
    vsize = 2 if cfg["dtype"] == ivr.IVR_F16 else 4
    bytes_alg = nvox * (vsize + 16)
    achieved = bytes_alg / (ms * 1e-3)
    threads = cpu_threads()
    out = {
        "metric": "Gvoxels/s, gradient field 512^3 fp32", "value": round(nvox / ms / 1e6, 3),
        "unit": "Gvoxels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "gradient field of " + cfg["name"], "voxels": nvox,
                   "bit_identical_to_oracle": same},
        "roofline": {"bound": "hbm", "kernel": "k_gradient<float>", "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": measured_traffic("gradient field of " + cfg["name"], "gradient", False)},
        "cpu_baseline": {"value": round(nvox / cpu_s / 1e9, 3), "unit": "Gvoxels/s", "cores": threads,
                         "kind": "port", "host": host_info(), "sample": "the whole %dx%dx%d volume, one timed run" % (nx, ny, nz)},
    }
    print(json.dumps(out), flush=True)
    vol.close()
    ctx.close()


def measured_traffic(workload, comp_name, skip):
    """HBM bytes per launch from the committed PMC summary, or None"""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    for e in json.load(open(path)).get("entries", []):
        if e["workload"] == workload and e["composite"] == comp_name and e["skip"] == skip:
            return int(e["fetch_kib"] * e["fetch_scale"] * 1024 + e["write_kib"] * e["write_scale"] * 1024)
    return None


def copy_calibration(mib):
    """hipMemcpy device-to-device of a known size, 5 times: the copy kernel
    reads and writes exactly mib MiB per dispatch"""
    import ctypes as C
    import ivr_amd  # noqa: F401  (loads the HIP runtime the library uses)
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(5):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 5}))


def host_info():
    """the host the CPU legs ran on: CPU model, CPUs in this process's
    affinity mask, OMP_NUM_THREADS"""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": model, "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads():
    """threads of the CPU legs: the CPUs this process may run on, capped by
    OMP_NUM_THREADS where the environment sets it (the box's CPU share)"""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def _oracle_frame(oracle, host, st, lut, comp, brick, rows, threads, f64=False):
    """the oracle's render of rows [r0, r1) of the same workload: one pass
    per ray, fused or sort-last (per-brick runs composited far to near); the
    FP64 sort-last form is the per-brick tiles composited in visibility order"""
    import numpy as np
    import ivr_amd as ivr
    v = oracle.view_of(st)
    if comp == ivr.FUSED:
        return oracle.render(host, v, lut, oracle.params(), rows=rows, nthreads=threads, f64=f64)
    if not f64:
        return oracle.render(host, v, lut, oracle.params(brick=brick, sort_last=True), rows=rows,
                             nthreads=threads)
    import ctypes as C
    cam = np.zeros(23, np.float32)
    rect = (C.c_int * 4)()
    bmin, bmax = oracle.box(host.shape[2], host.shape[1], host.shape[0])
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    oracle.lib.ivro_camera(fp(bmin), fp(bmax), C.byref(v), fp(cam), rect)
    dims = (host.shape[2], host.shape[1], host.shape[0])
    nb = [(dims[r] - 2) // brick + 1 for r in range(3)]
    eb = [int(np.floor((cam[r] - bmin[r]) * (dims[r] / (bmax[r] - bmin[r])) / brick)) for r in range(3)]
    ids = sorted([(x, y, z) for z in range(nb[2]) for y in range(nb[1]) for x in range(nb[0])],
                 key=lambda b: (sum(abs(b[r] - eb[r]) for r in range(3)), b[0] + nb[0] * (b[1] + nb[1] * b[2])))
    d, s = None, 0
    tiles = []
    for b in ids:
        t, k = oracle.render(host, v, lut, oracle.params(brick=brick, brick_id=b), rows=rows, nthreads=threads,
                             f64=True)
        tiles.append(t)
        s += k
    d = np.zeros_like(tiles[0])  # compositetiles in double: far to near, d = s + (1 - s.a) d
    for t in reversed(tiles):
        d = t + (1.0 - t[..., 3:4]) * d
    return d, s


def cpu_baseline(vol, st, lut, gpu_img, W, H, comp, brick, target_s):
    """The oracle (the CPU restatement of the same workload: same samples,
    same float operations, one pass per ray) on the host cores.  A timed run
    renders the whole frame `reps` times, or a band of rows through the image
    centre when one frame takes longer than target_s / 5; median of 5 timed
    runs after a warm-up."""
    import numpy as np
    import oracle
    host = vol.download()
    threads = cpu_threads()
    mid = H // 2
    t = time.perf_counter()
    _oracle_frame(oracle, host, st, lut, comp, brick, (mid - 4, mid + 4), threads)
    frame_s = (time.perf_counter() - t) * H / 8.0
    per_run = target_s / 5.0
    if frame_s <= per_run:
        r0, r1, reps = 0, H, max(1, int(per_run / max(frame_s, 1e-4)))
    else:
        rows = int(min(H, max(8, H * per_run / frame_s)))
        r0, r1, reps = max(0, mid - rows // 2), min(H, mid - rows // 2 + rows), 1
    log("cpu baseline: %d threads, rows %d..%d, %d frame(s) per timed run" % (threads, r0, r1, reps))
    _oracle_frame(oracle, host, st, lut, comp, brick, (r0, r1), threads)  # warm-up
    times = []
    for _ in range(5):
        t = time.perf_counter()
        for _ in range(reps):
            img32, n = _oracle_frame(oracle, host, st, lut, comp, brick, (r0, r1), threads)
        times.append((time.perf_counter() - t) / reps)
    dt = float(np.median(times))
    ref8 = oracle.quantize(img32[r0:r1])
    d = int(np.abs(ref8.astype(int) - gpu_img[r0:r1].astype(int)).max())
    # FP64 oracle on a narrower band: the error of the fp32 GPU result
    f0, f1 = mid - 8, mid + 8
    i64, _ = _oracle_frame(oracle, host, st, lut, comp, brick, (f0, f1), threads, f64=True)
    d64 = int(np.abs(oracle.quantize(i64[f0:f1]).astype(int) - gpu_img[f0:f1].astype(int)).max())
    log("cpu baseline: %.3fs per pass over %d samples; max |gpu-cpu| on those rows = %d/255, vs fp64 %d/255"
        % (dt, n, d, d64))
    return {
        "value": round(n / dt / 1e9, 4),
        "unit": "Gsamples/s",
        "cores": threads,
        "kind": "port",
        "host": host_info(),
        "build": "gcc -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp, OpenMP static schedule over rows",
        "sample": "rows %d-%d of the %dx%d frame, %d samples, %s, %d frame(s) per timed run, median of 5 runs "
                  "(%.3fs per frame) after 1 warm-up; max |gpu-cpu| %d/255; max |gpu-fp64 oracle| %d/255 on rows "
                  "%d-%d" % (r0, r1, W, H, n, "one pass per ray" if comp == 0 else
                             "one pass per ray, per-brick runs combined far to near", reps, dt, d, d64, f0, f1),
    }


class stdout_to_stderr:
    """route file descriptor 1 to 2 (the reference prints progress to stdout)"""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        os.dup2(self.saved, 1)
        os.close(self.saved)


def reference_cpu(ctx, vol, st, lut, W, H):
    """The reference's own ray.c/transferfuncs.c (compiled from its sources)
    rendering one frame of the same volume (values x255), camera and LUT with
    CHEM, its OpenMP loop on the CPU legs' thread count and on one core;
    against the GPU compat mode on identical input, whose image must be
    bit-identical."""
    import ctypes as C
    import numpy as np
    import ivr_amd as ivr
    import oracle
    ref = oracle.ref_lib()
    if ref is None:
        return None
    if max(vol.info()["dims"]) > 512:
        # getsamplesonray's last sample may pass the exit face by half a
        # 0.001-step (ray.c:253-261), more than a voxel beyond 512 per axis, and
        # getlinearsample has no lower index clamp (ray.c:126-135): an exit
        # through a low face then reads a slice before the volume buffer
        log("reference ray.c: not run on a %s volume (it reads before its buffer there)"
            % "x".join(str(d) for d in vol.info()["dims"]))
        return None
    host = vol.download() * np.float32(255.0)
    nz, ny, nx = host.shape
    gv = ctx.volume(nx, ny, nz, ivr.IVR_F32, brick=vol.info()["brick"])
    gv.upload(host)
    tf = ctx.tf(lut)
    fr = ctx.frame(W, H, keep_rgba32=True)
    d = ivr.compat_desc(st, xfer=ivr.CHEM, maxlutindx=255)
    ctx.render(gv, tf, d, fr)
    ctx.sync()
    ctx.record(0)
    for _ in range(5):
        ctx.render(gv, tf, d, fr)
    ctx.record(1)
    ctx.sync()
    g_ms = ctx.elapsed(0, 1) / 5
    S = fr.stats()["samples"]
    _, g32 = fr.download(rgba32=True)
    fr.close()
    tf.close()
    gv.close()
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    fd, path = tempfile.mkstemp(suffix=".raw")
    os.close(fd)
    try:
        host.tofile(path)
        with stdout_to_stderr():
            rc = ref.ivrref_init(path.encode(), nx, ny, nz, -1, 3, 8)
    finally:
        os.unlink(path)
    del host
    if rc != 0:
        return None
    lutc = np.ascontiguousarray(lut, np.float32)
    ref.ivrref_set_lut(f(lutc), lutc.shape[0], 0, 255)
    ref.ivrref_set_xfer(ivr.CHEM)
    mvm = np.array(st.mvm, np.float32)
    pm = np.array(st.pm, np.float32)
    vv = np.array(st.vv, np.float32)
    nc = np.array(st.newconst, np.float32)
    out = np.zeros((H, W, 4), np.float32)
    rect = (C.c_int * 4)()
    threads = cpu_threads()
    timed = {}
    for nt in (threads, 1):  # the reference's OpenMP pixel loop (main.c:983) on the box's share, then one core
        used = ref.ivrref_set_threads(nt)
        with stdout_to_stderr():
            t = time.perf_counter()
            ref.ivrref_render(f(mvm), f(pm), f(vv), f(nc), C.c_float(st.alpha), W, H, f(out), rect)
            timed[used] = time.perf_counter() - t
        if nt == threads:
            same = bool(np.array_equal(out.view(np.uint32), g32.view(np.uint32)))
    c_s = timed[max(timed)]
    log("reference ray.c: %s for %d samples; GPU compat %.3f ms; images identical: %s"
        % (", ".join("%.2fs on %d threads" % (v, k) for k, v in sorted(timed.items())), S, g_ms, same))
    return {
        "value": round(S / c_s / 1e9, 5),
        "unit": "Gsamples/s",
        "cores": max(timed),
        "kind": "reference",
        "host": host_info(),
        "one_core_gsamples_s": round(S / timed[min(timed)] / 1e9, 5) if min(timed) == 1 else None,
        "sample": "one %dx%d frame of the reference compat path (ray.c: 256-sample cap, 0.001 step, CHEM "
                  "back-to-front; its own OpenMP loop over rows, built as its makefile does) over the same "
                  "volume x255, camera and LUT; %d samples" % (W, H, S),
        "gpu_same_path_gsamples_s": round(S / g_ms / 1e6, 3),
        "gpu_same_path_ms": round(g_ms, 4),
        "gpu_over_reference": round((S / g_ms / 1e6) / (S / c_s / 1e9), 1),
        "images_identical": same,
    }


if __name__ == "__main__":
    main()
