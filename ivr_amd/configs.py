"""Workload presets of the five benchmark configurations.

Seeds: volume 1000+c, transfer function 2000+c, camera 3000+c (c = config
number).  The transfer function has 5 piecewise-linear control points; value 0
is transparent and the other control points have opacity U[0, ALPHA_MAX] so a
ray integrates hundreds of samples before early termination.
"""
import ivr_amd as ivr

ALPHA_MAX = 0.05
STEP_VOX = 0.5
ERT = 0.99

WORKLOADS = {
    1: dict(name="C1 64^3 fp32 blobs -> 256^2", dims=(64, 64, 64), dtype=ivr.IVR_F32, brick=64,
            image=(256, 256), skip=False, composite=ivr.FUSED,
            gen=dict(kind=ivr.GEN_BLOBS, nblobs=8)),
    2: dict(name="C2 512^3 fp32 Perlin+blobs (40% zero), 2x2x2 bricks of 256^3, sort-last -> 1024^2",
            dims=(512, 512, 512), dtype=ivr.IVR_F32, brick=256, image=(1024, 1024), skip=False,
            composite=ivr.SORT_LAST,
            gen=dict(kind=ivr.GEN_NOISE_BLOBS, nblobs=32, empty_frac=0.4)),
    3: dict(name="C3 1024^3 fp16 Perlin+blobs (70% zero), 4x4x4 bricks of 256^3, sort-last -> 1920x1080",
            dims=(1024, 1024, 1024), dtype=ivr.IVR_F16, brick=256, image=(1920, 1080), skip=True,
            composite=ivr.SORT_LAST, gen=dict(kind=ivr.GEN_NOISE_BLOBS, nblobs=32, empty_frac=0.7)),
}


def gen_params(c, dims=None):
    cfg = WORKLOADS[c]
    g = dict(cfg["gen"])
    return ivr.gen_params(seed=1000 + c, **g)


def tf_lut(c, jitter=0.0, jitter_seed=0):
    return ivr.tf_build(2000 + c, ALPHA_MAX, jitter, jitter_seed)


def state(c, width=None, height=None, seed=None):
    w, h = WORKLOADS[c]["image"]
    return ivr.orbit_state(3000 + c if seed is None else seed, width or w, height or h)
