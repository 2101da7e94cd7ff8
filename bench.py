"""Benchmark of the MI355X render path on one GPU.

Workload (the headline configuration): a 512^3 fp32 volume (4-octave Perlin
fBm + 32 Gaussian blobs, zeroed below its 40th percentile) stored as 2x2x2
bricks of 256^3 with a one-voxel ghost layer, rendered to a 1024x1024 RGBA8
frame with a 256-entry transfer function, 0.5-voxel steps and early ray
termination at alpha >= 0.99 (config mode).  The volume, LUT and frame are
resident in HBM before timing; a step is one full frame.

value = Gsamples/s = composited samples per frame / device time per frame.
roofline: HBM bound; algorithmic bytes per frame = S * 32 B (8 fp32 voxels
per trilinear sample, no reuse) + 4 B per output pixel, divided by the
average device time of a frame (HIP events on the library's stream), against
8 TB/s.  cpu_baseline: the oracle's config-mode CPU restatement (same
samples, same arithmetic) timed on host cores over a band of rows.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]
"""
import argparse
import json
import os
import sys
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
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
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

    import numpy as np
    import ivr_amd as ivr
    from ivr_amd import configs

    c = args.config
    cfg = configs.CONFIGS[c]
    W, H = cfg["image"]
    skip = cfg["skip"] if args.skip < 0 else bool(args.skip)
    ctx = ivr.Context(0)
    log("generating", cfg["name"])
    vol = ctx.volume(*cfg["dims"], cfg["dtype"], brick=cfg["brick"])
    t0 = time.time()
    vol.generate(configs.gen_params(c))
    log("generated in %.2fs" % (time.time() - t0))
    lut = configs.tf_lut(c)
    tf = ctx.tf(lut)
    st = configs.state(c)
    desc = ivr.config_desc(st, skip=skip)
    fr = ctx.frame(W, H)

    for _ in range(args.warmup):
        ctx.render(vol, tf, desc, fr)
    ctx.sync()
    stats = fr.stats()
    S = stats["samples"]
    log("warmup done: samples/frame %d, fetched %d, last frame %.3f ms" % (S, stats["visited"], stats["render_ms"]))

    ctx.sync()
    t_wall = time.perf_counter()
    ctx.record(0)
    for _ in range(args.steps):
        ctx.render(vol, tf, desc, fr)
    ctx.record(1)
    ctx.sync()
    t_wall = time.perf_counter() - t_wall
    dev_ms = ctx.elapsed(0, 1) / args.steps
    ms_step = max(dev_ms, 1e-9)
    gsps = S / (ms_step * 1e-3) / 1e9
    vox_bytes = 2 if cfg["dtype"] == ivr.IVR_F16 else 4
    fetched = stats["visited"]
    bytes_alg = fetched * 8 * vox_bytes + 4 * W * H
    achieved = bytes_alg / (ms_step * 1e-3)
    log("timed %d frames: %.3f ms/frame device, %.3f ms/frame wall, %.1f Gsamples/s"
        % (args.steps, dev_ms, t_wall * 1e3 / args.steps, gsps))

    img = fr.download()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(vol, st, lut, img, W, H, args.cpu_seconds)

    out = {
        "metric": "Gsamples/s, ms/frame 512³→1024²; surrogate fps @PSNR≥40dB; p99 latency",
        "value": round(gsps, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if cfg["dtype"] == ivr.IVR_F32 else "f16 storage, f32 compute",
        "data": "synthetic (seeded Perlin fBm + Gaussian blobs, generated on the GPU)",
        "config": {
            "workload": cfg["name"],
            "image": "%dx%d" % (W, H),
            "samples_per_frame": int(S),
            "fetched_per_frame": int(fetched),
            "step_voxels": configs.STEP_VOX,
            "ert": configs.ERT,
            "macrocell_skip": bool(skip),
            "frames_per_s": round(1000.0 / ms_step, 2),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved / 1e9, 1),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4),
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    if cpu:
        out["config"]["gpu_over_cpu"] = round(gsps / cpu["value"], 1) if cpu["value"] > 0 else None
    print(json.dumps(out), flush=True)
    fr.close()
    tf.close()
    vol.close()
    ctx.close()


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


def cpu_baseline(vol, st, lut, gpu_img, W, H, target_s):
    """Oracle config mode on a band of rows through the image centre."""
    import numpy as np
    import oracle
    host = vol.download()
    threads = min(16, os.cpu_count() or 1)
    v = oracle.view_of(st)
    p = oracle.params()
    mid = H // 2
    r0, r1 = mid - 4, mid + 4
    t = time.perf_counter()
    _, n = oracle.render(host, v, lut, p, rows=(r0, r1), nthreads=threads)
    dt = time.perf_counter() - t
    rows = int(min(H, max(8, 8 * target_s / max(dt, 1e-3))))
    r0, r1 = max(0, mid - rows // 2), min(H, mid - rows // 2 + rows)
    log("cpu baseline: %d threads, rows %d..%d" % (threads, r0, r1))
    oracle.render(host, v, lut, p, rows=(r0, r1), nthreads=threads)  # warm-up
    times = []
    while len(times) < 5 and sum(times) < target_s:
        t = time.perf_counter()
        img32, n = oracle.render(host, v, lut, p, rows=(r0, r1), nthreads=threads)
        times.append(time.perf_counter() - t)
    dt = float(np.median(times))
    ref8 = oracle.quantize(img32[r0:r1])
    d = int(np.abs(ref8.astype(int) - gpu_img[r0:r1].astype(int)).max())
    log("cpu baseline: %.2fs for %d samples; max |gpu-cpu| on those rows = %d/255" % (dt, n, d))
    return {
        "value": round(n / dt / 1e9, 4),
        "unit": "Gsamples/s",
        "cores": threads,
        "kind": "port",
        "sample": "rows %d-%d of the %dx%d frame, %d samples, median of %d timed runs (%.2fs each) after "
                  "1 warm-up; max |gpu-cpu| %d/255" % (r0, r1, W, H, n, len(times), dt, d),
    }


if __name__ == "__main__":
    main()
