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

The default run also carries the rest of the metric in the same line:
"survey_tf" (the headline frame with control-point RGBA U[0,1]),
"surrogate" (frames/s at 1920x1080 with per-brick fallback below 40 dB, for
the surrogate of the C2 volume and of the 1024^3 fp16 volume, each trained
on the GPU for --train-seconds) and "serving" (p50/p99 latency and frames/s
of the steering mix at 50-400 req/s, plain and with the 1024^3 resident
served from its surrogate).  --config 4, --config 5 and --gradient (the
central-difference gradient field of the C2 volume) print their own lines.

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
    ap.add_argument("--train-seconds", type=float, default=None,
                    help="surrogate training time budget per surrogate (default: 20 s in the headline run, "
                         "60 s with --config 4/5)")
    ap.add_argument("--headline-only", action="store_true",
                    help="the C2 frame rate alone, without the survey-TF, surrogate and serving parts")
    ap.add_argument("--sg-volume", type=int, default=2, choices=[2, 3],
                    help="--config 4: train on the C2 (512^3 fp32) or the C3 (1024^3 fp16) volume")
    ap.add_argument("--sg-group", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="--config 4: one hash grid per 2^g volume bricks per axis (g > 0: frozen compact "
                         "surrogate, compared with the plain field at the same bytes)")
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

    if args.train_seconds is None:
        args.train_seconds = 60.0 if args.config in (4, 5) else 20.0
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
            "traffic_frac": round(traffic / (k_ms * 1e-3) / HBM_PEAK, 4) if traffic else None,
            "compulsory_bytes": compulsory_bytes(cfg, W, H),
            "l2_hit": measured_l2_hit(cfg["name"], comp_name, skip),
            "note": "achieved counts 8 voxels per classified sample with no reuse (plus the partial "
                    "planes); neighbouring samples and rays share footprint voxels in L1/L2, so this "
                    "rate can pass the HBM peak; traffic is the measured HBM bytes per launch and "
                    "traffic_frac their rate against the peak; compulsory_bytes is the floor of "
                    "ghosted brick bytes touched plus the image; l2_hit from the committed counters",
        },
        "cpu_baseline": cpu,
        "reference_cpu": ref,
    }
    if cpu and cpu["value"] > 0:
        out["config"]["gpu_over_cpu"] = round(gsps / cpu["value"], 1)
    if c == 2 and not args.headline_only:
        # the rest of the metric: the same frame under the survey's control-point opacities,
        # surrogate frames/s at PSNR >= 40 dB (512^3 and 1024^3), steering p50/p99 and frames/s
        out["survey_tf"] = survey_tf_measure(ctx, vol, st, comp, W, H, cfg["brick"], args.steps)
        fr.close()
        tf.close()
        sur = {}
        sur["c4_512"], _ = surrogate_measure(ctx, vol, 2, args.train_seconds, args.steps,
                                             0.0 if args.no_cpu else 5.0)
        cfg3 = configs.WORKLOADS[3]
        v3 = ctx.volume(*cfg3["dims"], cfg3["dtype"], brick=cfg3["brick"])
        v3.generate(configs.gen_params(3))
        out["c3"] = c3_measure(ctx, v3, args.steps)
        # the 1024^3 surrogate in grids of 512^3 (2x2x2 volume bricks), frozen: 1/8 of the fp16 field
        sur["vol_1024"], sg3 = surrogate_measure(ctx, v3, 3, args.train_seconds, args.steps, keep=True, group=1)
        out["surrogate"] = sur
        vols = serving_volumes(ctx, c2=vol, c3=v3)
        n_req = 200
        plain_rows, _ = serving_sweep(ctx, vols, n_req, (50.0, 100.0, 200.0, 400.0, 1e9))
        sg_rows, _ = serving_sweep(ctx, vols, n_req, (50.0, 100.0, 1e9), sg3, "surrogate")
        out["serving"] = {
            "workload": "C5: 256^3 fp32, 512^3 fp32, 1024^3 fp16 resident; 16 slots; 64 GiB budget; %d seeded "
                        "requests per rate, resolutions {256^2, 512^2, 1024^2, 1920x1080}" % n_req,
            "plain": plain_rows, "capacity_frames_per_s": plain_rows[-1]["frames_per_s"],
            "p99_ms_at_400": plain_rows[3]["p99_ms"],
            "with_1024_surrogate": sg_rows, "surrogate_capacity_frames_per_s": sg_rows[-1]["frames_per_s"]}
        sg3.close()
        vols[0].close()
        v3.close()
    else:
        fr.close()
        tf.close()
    print(json.dumps(out), flush=True)
    vol.close()
    ctx.close()


def c3_measure(ctx, vol, steps):
    """The C3 frame on a resident C3 volume (1024^3 fp16, 4x4x4 bricks of
    256^3, macrocell skipping, sort-last, 1920x1080): ms/frame, Gsamples/s,
    the k_runs roofline, and the GPU image against the oracle over a band of
    rows through the centre"""
    import numpy as np
    import ivr_amd as ivr
    import oracle
    from ivr_amd import configs
    cfg = configs.WORKLOADS[3]
    W, H = cfg["image"]
    lut = configs.tf_lut(3)
    tf = ctx.tf(lut)
    st = configs.state(3)
    d = ivr.config_desc(st, skip=True, composite=ivr.SORT_LAST)
    fr = ctx.frame(W, H)
    ms = _frame_ms(ctx, lambda: ctx.render(vol, tf, d, fr), steps)
    stats = fr.stats()
    S = int(stats["visited"])
    img = fr.download()
    k_ms, c_ms = kernel_ms(ctx, lambda: ctx.render(vol, tf, d, fr), fr, steps)
    # 8 fp16 voxels per classified sample, the partials the rays have, a run count per pixel
    bytes_alg = S * 16 + int(stats["partials"]) * 16 + 4 * W * H
    r0, r1 = H // 2 - 8, H // 2 + 8
    ref32, _ = _oracle_frame(oracle, vol.download(), st, lut, ivr.SORT_LAST, cfg["brick"], (r0, r1), cpu_threads())
    dmax = int(np.abs(oracle.quantize(ref32[r0:r1]).astype(int) - img[r0:r1].astype(int)).max())
    fr.close()
    tf.close()
    log("C3: %.4f ms/frame, %d samples classified, %.1f Gsamples/s, k_runs %.4f ms; max |gpu - oracle| %d/255 "
        "on rows %d-%d" % (ms, S, S / ms / 1e6, k_ms, dmax, r0, r1))
    return {"workload": cfg["name"], "ms_per_frame": round(ms, 4), "frames_per_s": round(1000.0 / ms, 2),
            "samples_per_frame": S, "samples_incl_skipped": int(stats["samples"]),
            "gsamples_s": round(S / ms / 1e6, 2),
            "roofline": {"bound": "hbm", "kernel": "k_runs<_Float16, true, true>", "kernel_ms": round(k_ms, 4),
                         "combine_ms": round(c_ms, 4), "algorithmic_bytes": bytes_alg,
                         "achieved": round(bytes_alg / (k_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK / 1e9,
                         "unit": "GB/s", "frac": round(bytes_alg / (k_ms * 1e-3) / HBM_PEAK, 4),
                         "traffic": measured_traffic(cfg["name"], "sort_last", True),
                         "compulsory_bytes": compulsory_bytes(cfg, W, H),
                         "l2_hit": measured_l2_hit(cfg["name"], "sort_last", True)},
            "max_abs_diff_vs_oracle": dmax, "oracle_rows": [r0, r1]}


def survey_tf_measure(ctx, vol, st, comp, W, H, brick, steps):
    """The headline frame with control-point RGBA drawn U[0,1] (the value-0
    point transparent, so the zeroed part stays empty space) instead of
    opacities capped at 0.05: ms/frame, Gsamples/s, and the GPU image against
    the oracle over a band of rows through the centre"""
    import numpy as np
    import ivr_amd as ivr
    import oracle
    lut = ivr.tf_build(2002, 1.0)
    tf = ctx.tf(lut)
    fr = ctx.frame(W, H)
    d = ivr.config_desc(st, composite=comp)
    ctx.render(vol, tf, d, fr)
    ctx.sync()
    ctx.record(0)
    for _ in range(steps):
        ctx.render(vol, tf, d, fr)
    ctx.record(1)
    ctx.sync()
    ms = ctx.elapsed(0, 1) / steps
    S = fr.stats()["visited"]
    img = fr.download()
    r0, r1 = H // 2 - 16, H // 2 + 16
    ref32, s_band = _oracle_frame(oracle, vol.download(), st, lut, comp, brick, (r0, r1), cpu_threads())
    dmax = int(np.abs(oracle.quantize(ref32[r0:r1]).astype(int) - img[r0:r1].astype(int)).max())
    fr.close()
    tf.close()
    log("survey TF: %.4f ms/frame, %d samples, %.1f Gsamples/s; max |gpu - oracle| %d/255 on rows %d-%d"
        % (ms, S, S / ms / 1e6, dmax, r0, r1))
    return {"tf": "5 control points, RGBA U[0,1] (value 0 transparent)", "ms_per_frame": round(ms, 4),
            "samples_per_frame": int(S), "gsamples_s": round(S / ms / 1e6, 2),
            "max_abs_diff_vs_oracle": dmax, "oracle_rows": [r0, r1]}


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


def train_surrogate(ctx, vol, seconds, batch=1 << 24, chunk=5, group=0):
    """Adam steps of `batch` samples in chunks until `seconds` of training"""
    sg = ctx.surrogate(vol, log2_group=group)
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


def lowres_field(vol, f):
    """the plain method at a surrogate's footprint: vol's values at every f-th
    grid point per axis, rounded to fp16 (1/f^3 of the fp16 field's bytes), and
    the full grid reconstructed from them by trilinear interpolation (host
    array, float32)"""
    import numpy as np
    a = vol.download()
    shape = a.shape
    g = a[::f, ::f, ::f].astype(np.float16).astype(np.float32)
    del a
    for axis in range(3):
        n, K = shape[axis], g.shape[axis]
        i = np.arange(n)
        k = np.minimum(i // f, K - 1)
        k1 = np.minimum(k + 1, K - 1)
        bshape = [1, 1, 1]
        bshape[axis] = n
        w = ((i - k * f) / f).astype(np.float32).reshape(bshape)
        out = np.take(g, k, axis)
        out *= 1.0 - w
        hi = np.take(g, k1, axis)
        hi *= w
        out += hi
        del hi
        g = out
    return g


def surrogate_measure(ctx, vol, c, train_seconds, steps, cpu_seconds=0.0, keep=False, group=0):
    """A surrogate of `vol` (config c's volume: 2 = the 512^3 fp32 C2 volume,
    3 = the 1024^3 fp16 C3 volume): hash grids (16 levels x 2^19 x 2
    features) per brick of 2^group volume bricks per axis + the BF16 MFMA MLP
    trained on the GPU with 2^24 samples per Adam step for a time budget, then
    1920x1080 frames from it against the plain render of the same camera:
    PSNR/SSIM, frames/s, per-brick fallback below 40 dB.  With group > 0 the
    surrogate is frozen to its half-precision tables and rendered without the
    plain volume (unless a brick falls back), its resident bytes are reported
    beside the field's, and the plain method at the same bytes (the field at
    1/2^group resolution in fp16, trilinearly reconstructed) is rendered for
    comparison.  Returns (dict, surrogate or None)."""
    import numpy as np
    import ivr_amd as ivr
    from ivr_amd import configs
    W, H = 1920, 1080
    tf = ctx.tf(configs.tf_lut(c))
    desc = ivr.config_desc(ivr.orbit_state(3004, W, H), skip=True)  # camera seed of C4; skipping on
    fr = ctx.frame(W, H)
    plain_ms = _frame_ms(ctx, lambda: ctx.render(vol, tf, desc, fr), steps)
    plain = fr.download()
    batch = 1 << 24
    b0 = ctx.device_bytes()
    sg, losses, nsteps, t_train = train_surrogate(ctx, vol, train_seconds, batch, group=group)
    train_bytes = ctx.device_bytes() - b0
    step_ms = t_train * 1e3 / nsteps
    sg.set_fallback(np.zeros(sg.nbricks, np.uint8))
    sg_ms = _frame_ms(ctx, lambda: sg.render(vol, tf, desc, fr), steps)
    sg_stats = fr.stats()
    img = fr.download()
    p_sg, s_sg = ivr.psnr(img, plain), ivr.ssim(img, plain)
    bpsnr, nfall = sg.evaluate(vol, tf, desc, threshold_db=40.0)
    cpu_sg = cpu_surrogate_baseline(sg, cpu_seconds) if cpu_seconds > 0 else None
    compact = None
    rvol = vol
    if group > 0:
        sg.freeze()
        rvol = vol if nfall else None  # the plain volume only for bricks that fall back
        compact = {"log2_group": group, "surrogate_bricks": sg.nbricks,
                   "resident_bytes_training": int(train_bytes),
                   "resident_bytes_frozen": int(ctx.device_bytes() - b0),
                   "field_bytes": int(vol.info()["bytes"]), "rendered_without_plain_volume": rvol is None}
    fb_ms = _frame_ms(ctx, lambda: sg.render(rvol, tf, desc, fr), steps)
    img_fb = fr.download()
    p_fb, s_fb = ivr.psnr(img_fb, plain), ivr.ssim(img_fb, plain)
    # the network kernel's own duration (HIP events around k_render_sg on the library's stream)
    k_ms, _ = kernel_ms(ctx, lambda: sg.render(rvol, tf, desc, fr), fr, steps)
    S = int(fr.stats()["visited"])
    if compact:
        f = 1 << group
        low = ctx.volume(*vol.shape[::-1], vol.info()["dtype"], brick=vol.info()["brick"])
        low.upload(lowres_field(vol, f))
        ctx.render(low, tf, desc, fr)
        img_low = fr.download()
        low.close()
        compact["frozen_over_field"] = round(compact["resident_bytes_frozen"] / compact["field_bytes"], 4)
        compact["equal_bytes_plain"] = {
            "method": "field at every %d-th voxel per axis in fp16 (%d bytes), trilinear reconstruction" % (
                f, 2 * int(np.prod([(n + f - 1) // f for n in vol.shape]))),
            "psnr_db": round(ivr.psnr(img_low, plain), 2), "ssim": round(ivr.ssim(img_low, plain), 4)}
        log("compact surrogate: %d bytes frozen (%.4f of the field), PSNR %.2f dB; plain field at 1/%d "
            "resolution: %.2f dB" % (compact["resident_bytes_frozen"], compact["frozen_over_field"], p_fb, f,
                                     compact["equal_bytes_plain"]["psnr_db"]))
    log("surrogate %s: train %.1f ms/step (%d steps), render %.3f ms (plain %.3f); PSNR %.2f dB SSIM %.4f; "
        "%d/%d bricks below 40 dB -> with fallback PSNR %.2f dB, %.3f ms"
        % ("C2" if c == 2 else "1024^3", step_ms, nsteps, sg_ms, plain_ms, p_sg, s_sg, nfall, sg.nbricks, p_fb,
           fb_ms))
    flops = 2 * (32 * 64 + 64 * 64 + 64)
    g_bytes = S * 16 * 8 * 4  # per network sample: 16 levels x 8 corners x 2 half-precision features
    workload = "surrogate of %s: %d bricks%s x 16 levels x 2^19 x 2 features, MLP 32-64-64-1, 1920x1080" \
        % ("C2 (512^3 fp32)" if c == 2 else "C3 (1024^3 fp16)", sg.nbricks,
           " of %d^3" % (vol.info()["brick"] << group) if group else "")
    out = {
        "workload": workload,
        "fps": round(1000.0 / fb_ms, 2), "ms_per_frame": round(fb_ms, 3),
        "psnr_db": round(p_fb, 2), "ssim": round(s_fb, 4),
        "fallback_bricks": int(nfall), "bricks": sg.nbricks,
        "brick_psnr_db": [round(float(x), 2) for x in bpsnr],
        "no_fallback": {"ms_per_frame": round(sg_ms, 3), "psnr_db": round(p_sg, 2), "ssim": round(s_sg, 4)},
        "network_samples_per_frame": int(sg_stats["visited"]),
        "network_lane_utilisation": round(sg_stats["visited"] / max(1, sg_stats["evaluations"]), 4),
        "plain_ms_per_frame": round(plain_ms, 3), "plain_fps": round(1000.0 / plain_ms, 2),
        "train": {"seconds": round(t_train, 1), "steps": nsteps, "samples_per_step": batch,
                  "ms_per_step": round(step_ms, 2),
                  "mlp_tflops": round(batch * 3 * flops / (step_ms * 1e-3) / 1e12, 2),
                  "loss_first": losses[0], "loss_last": losses[-1]},
        "roofline": {"bound": "hbm", "kernel": "k_render_sg<true>", "kernel_ms": round(k_ms, 4),
                     "achieved": round(g_bytes / (k_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(g_bytes / (k_ms * 1e-3) / HBM_PEAK, 4),
                     "traffic": measured_traffic("C4 surrogate of %s: %d bricks x 16 levels x 2^19 x 2 features, "
                                                 "MLP 32-64-64-1" % ("C2" if c == 2 else "C3", sg.nbricks),
                                                 "surrogate", True),
                     "mfma_tflops": round(S * flops / (k_ms * 1e-3) / 1e12, 2), "mfma_peak_tflops": 2500.0,
                     "mfma_frac": round(S * flops / (k_ms * 1e-3) / 2.5e15, 5)},
    }
    if compact:
        out["compact"] = compact
    if cpu_sg:
        out["cpu_baseline"] = cpu_sg
    fr.close()
    tf.close()
    if not keep:
        sg.close()
        sg = None
    return out, sg


def bench_surrogate(args):
    """C4 alone (--config 4): the surrogate of the C2 volume, or with
    --sg-volume 3 of the 1024^3 fp16 C3 volume"""
    import ivr_amd as ivr
    from ivr_amd import configs
    c = args.sg_volume
    cfg = configs.WORKLOADS[c]
    ctx = ivr.Context(0)
    vol = ctx.volume(*cfg["dims"], cfg["dtype"], brick=cfg["brick"])
    vol.generate(configs.gen_params(c))
    res, _ = surrogate_measure(ctx, vol, c, args.train_seconds, args.steps, 0.0 if args.no_cpu else args.cpu_seconds,
                               group=args.sg_group)
    out = {
        "metric": "surrogate frames/s at 1920x1080 with per-brick fallback below 40 dB",
        "value": res["fps"],
        "unit": "frames/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16 MLP on MFMA, fp32 tables and accumulation",
        "data": "synthetic %s volume, surrogate trained on the GPU" % ("C2" if c == 2 else "C3"),
        "config": res,
        "roofline": res.pop("roofline"),
        "cpu_baseline": res.pop("cpu_baseline", None),
    }
    print(json.dumps(out), flush=True)
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


def serving_volumes(ctx, c2=None, c3=None):
    """the three resident volumes of the steering mix: 256^3 fp32, the C2
    512^3 fp32 and the C3 1024^3 fp16 volume (generated unless given)"""
    import ivr_amd as ivr
    from ivr_amd import configs
    vols = []
    for (n, dt, c, given) in ((256, ivr.IVR_F32, 2, None), (512, ivr.IVR_F32, 2, c2), (1024, ivr.IVR_F16, 3, c3)):
        if given is not None:
            vols.append(given)
            continue
        v = ctx.volume(n, n, n, dt, brick=256)
        v.generate(configs.gen_params(c))
        vols.append(v)
    return vols


def serving_sweep(ctx, vols, n_requests, rates, sg=None, tag="plain"):
    """Poisson arrivals at each rate (req/s; >= 1e9: all at once, the
    capacity) into the continuous batcher with the three residents (the
    1024^3 one served from `sg` when given); each request a seeded orbit
    camera, a jittered TF, a resolution from {256^2, 512^2, 1024^2,
    1920x1080} and a volume; latency = arrival -> RGBA8 on the host"""
    import numpy as np
    from ivr_amd import serving
    bt = ctx.batcher(64 << 30, 16, 1920, 1080)
    bt.add_volume(vols[0])
    bt.add_volume(vols[1])
    if sg is None:
        bt.add_volume(vols[2])
    else:
        bt.add_surrogate(sg, vols[2])
    rows = []
    for k, rate in enumerate(rates):
        reqs = serving.make_requests(n_requests, len(vols), seed=4100 + k, tf_seed=2005)
        res = serving.run_poisson(bt, reqs, rate, seed=4000 + k)
        row = {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in res.items() if kk != "records"}
        row["rate"] = "burst" if rate >= 1e9 else rate
        lat3 = [(r["t_done"] - r["t_arrival"]) * 1e3 for r in res["records"].values()
                if r["status"] == 0 and reqs[r["id"] - 1]["volume"] == 2]
        if lat3:
            row["vol1024_p50_ms"] = round(float(np.percentile(lat3, 50)), 3)
            row["vol1024_p99_ms"] = round(float(np.percentile(lat3, 99)), 3)
        log("serving (%s): %s" % (tag, row))
        rows.append(row)
    st = bt.stats()
    bt.close()
    return rows, st


def bench_serving(args):
    """C5 alone (--config 5): the steering mix at 50..400 req/s and an
    all-at-once burst, optionally with the 1024^3 resident served from a
    surrogate trained for --train-seconds (--serve-surrogate)"""
    import ivr_amd as ivr
    ctx = ivr.Context(0)
    vols = serving_volumes(ctx)
    rows, st = serving_sweep(ctx, vols, args.requests, (50.0, 100.0, 200.0, 400.0, 1e9))
    sg_part = None
    if args.serve_surrogate:
        res, sg = surrogate_measure(ctx, vols[2], 3, args.train_seconds, args.steps, keep=True)
        srows, _ = serving_sweep(ctx, vols, args.requests, (50.0, 100.0, 200.0, 400.0, 1e9), sg, "surrogate")
        sg_part = {"surrogate": res, "rates": srows, "capacity_frames_per_s": srows[-1]["frames_per_s"]}
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
                   "capacity_frames_per_s": rows[-1]["frames_per_s"], "batcher": st,
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


def _traffic_entry(workload, comp_name, skip):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    for e in json.load(open(path)).get("entries", []):
        if e["workload"] == workload and e["composite"] == comp_name and e["skip"] == skip:
            return e
    return None


def measured_traffic(workload, comp_name, skip):
    """HBM bytes per launch from the committed PMC summary, or None"""
    e = _traffic_entry(workload, comp_name, skip)
    if e is None:
        return None
    return int(e["fetch_kib"] * e["fetch_scale"] * 1024 + e["write_kib"] * e["write_scale"] * 1024)


def measured_l2_hit(workload, comp_name, skip):
    """L2 hit rate of the kernel from the committed PMC summary, or None"""
    e = _traffic_entry(workload, comp_name, skip)
    return e.get("l2_hit") if e else None


def compulsory_bytes(cfg, W, H):
    """the compulsory-traffic floor of a frame: the ghosted bytes of the bricks
    it touches (the orbit cameras keep the whole volume in view, so every
    brick) plus the RGBA8 image"""
    import ivr_amd as ivr
    nb = 1
    for d in cfg["dims"]:
        nb *= -(-d // cfg["brick"])
    esz = 4 if cfg["dtype"] == ivr.IVR_F32 else 2
    return nb * (cfg["brick"] + 2) ** 3 * esz + 4 * W * H


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


def _load_refgpu():
    """the reference's own main.c/mathutil.c/utils.c linked against
    libivr_compat.so (oracle/_ref/libivrref_gpu.so), or None"""
    import ctypes as C
    path = os.path.join(ROOT, "oracle", "_ref", "libivrref_gpu.so")
    return C.CDLL(path) if os.path.exists(path) else None


def dropin_per_ray(refgpu, lut, st, W, H, want):
    """The unchanged reference host relinked against the drop-in: its own
    createlocalimage pixel loop (OpenMP over rows, main.c:983-1009) calling
    the drop-in raycast per pixel (calls of concurrent threads combined into
    shared launches), and the same host with the loop replaced by
    ivr_compat_createlocalimage (one launch); ms/frame of each and whether
    their images equal the reference's"""
    import ctypes as C
    import numpy as np
    import ivr_amd as ivr
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    refgpu.ivrref_set_lut(f(lut), lut.shape[0], 0, 255)
    refgpu.ivrref_set_xfer(ivr.CHEM)
    args = (f(np.array(st.mvm, np.float32)), f(np.array(st.pm, np.float32)), f(np.array(st.vv, np.float32)),
            f(np.array(st.newconst, np.float32)), C.c_float(st.alpha), W, H)
    res = {}
    for name, gpu_frame in (("per_pixel_raycast", 0), ("createlocalimage_one_launch", 1)):
        refgpu.ivrref_frame_on_gpu(gpu_frame)
        out = np.zeros((H, W, 4), np.float32)
        rect = (C.c_int * 4)()
        l0, r0 = C.c_ulonglong(), C.c_ulonglong()
        refgpu.ivr_compat_raycast_stats(C.byref(l0), C.byref(r0))
        with stdout_to_stderr():
            refgpu.ivrref_render(*args, f(out), rect)  # warm-up: volume and LUT mirrored to the device
            t = time.perf_counter()
            refgpu.ivrref_render(*args, f(out), rect)
            ms = (time.perf_counter() - t) * 1e3
        l1, r1 = C.c_ulonglong(), C.c_ulonglong()
        refgpu.ivr_compat_raycast_stats(C.byref(l1), C.byref(r1))
        e = {"ms_per_frame": round(ms, 2),
             "images_identical": bool(np.array_equal(out.view(np.uint32), want.view(np.uint32)))}
        if gpu_frame == 0:
            rays, launches = (r1.value - r0.value) // 2, (l1.value - l0.value) // 2
            e.update(rays=int(rays), launches=int(launches), rays_per_launch=round(rays / max(1, launches), 2))
        res[name] = e
        log("drop-in %s: %.2f ms/frame %s" % (name, ms, e))
    refgpu.ivrref_frame_on_gpu(0)
    return res


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
            refgpu = _load_refgpu()
            if refgpu is not None and refgpu.ivrref_init(path.encode(), nx, ny, nz, -1, 3, 8) != 0:
                refgpu = None
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
    ref.ivrref_set_threads(threads)  # OpenMP's thread count is shared by every library of the process
    c_s = timed[max(timed)]
    log("reference ray.c: %s for %d samples; GPU compat %.3f ms; images identical: %s"
        % (", ".join("%.2fs on %d threads" % (v, k) for k, v in sorted(timed.items())), S, g_ms, same))
    dropin = None if refgpu is None else dropin_per_ray(refgpu, lutc, st, W, H, out)
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
        "dropin_relinked_host": dropin,
    }


if __name__ == "__main__":
    main()
