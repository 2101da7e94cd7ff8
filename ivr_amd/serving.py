"""Steering-serving load driver: Poisson request arrivals into the
continuous batcher (ivr_batcher_*), latency measured from each request's
scheduled arrival to its RGBA8 frame on the host.

Each request draws (seeded): a resident volume uniformly, an orbit camera,
a transfer function whose control points are jittered by U[-0.05, 0.05] and
a resolution uniformly from RESOLUTIONS.
"""
import time

import numpy as np

import ivr_amd as ivr

RESOLUTIONS = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]


def make_requests(n, nvols, seed, tf_seed, resolutions=RESOLUTIONS, jitter=0.05, skip=True):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        w, h = resolutions[rng.integers(len(resolutions))]
        vol = int(rng.integers(nvols))
        st = ivr.orbit_state(int(rng.integers(1 << 31)), int(w), int(h))
        lut = ivr.tf_build(tf_seed, 0.05, jitter, int(rng.integers(1 << 31)))
        reqs.append(dict(id=k + 1, volume=vol, desc=ivr.config_desc(st, skip=skip), lut=lut))
    return reqs


def run_poisson(batcher, reqs, rate, seed):
    """submit reqs at Poisson arrival times of the given rate (req/s);
    returns latency percentiles (ms), frames/s and per-request records"""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    t0 = ivr.now() + 0.01
    arrivals = t0 + np.cumsum(gaps)
    done = {}
    k = 0
    while len(done) < len(reqs):
        now = ivr.now()
        while k < len(reqs) and arrivals[k] <= now:
            r = reqs[k]
            batcher.submit(r["id"], r["volume"], r["desc"], r["lut"])
            k += 1
        batcher.progress(False)
        for c in batcher.poll():
            done[c["id"]] = c
        if k < len(reqs) and not batcher.busy():
            time.sleep(max(0.0, min(arrivals[k] - ivr.now(), 0.002)))
    for c in done.values():
        c["t_arrival"] = float(arrivals[c["id"] - 1])
    t_end = max(c["t_done"] for c in done.values())
    lat = np.array([(c["t_done"] - arrivals[c["id"] - 1]) * 1e3 for c in done.values() if c["status"] == 0])
    ok = int((np.array([c["status"] for c in done.values()]) == 0).sum())
    return dict(rate=rate, requests=len(reqs), ok=ok, p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
                frames_per_s=ok / (t_end - t0), wall_s=t_end - t0, records=done)
