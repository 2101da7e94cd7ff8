"""Continuous batching of steering requests: every image is bit-identical
alone (ivr_render) and in a mixed batch of other sizes, cameras, transfer
functions and volumes; admission respects the HBM budget; requests that can
never fit are rejected, not queued forever."""
import numpy as np
import pytest

import ivr_amd as ivr
from ivr_amd import configs, serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vols(ctx):
    a = ctx.volume(96, 96, 96, ivr.IVR_F32, brick=32)
    a.generate(configs.gen_params(2))
    b = ctx.volume(64, 80, 72, ivr.IVR_F16, brick=32)
    b.upload(ivr.generate_host(configs.gen_params(1), 64, 80, 72))
    yield [a, b]
    a.close()
    b.close()


def test_alone_equals_mix(ctx, vols):
    sizes = [(64, 48), (128, 96), (200, 150), (256, 256), (33, 17)]
    reqs = serving.make_requests(14, 2, seed=11, tf_seed=2005, resolutions=sizes)
    for k, r in enumerate(reqs):
        r["desc"].skip = k % 3 == 0
    bt = ctx.batcher(1 << 34, 4, 256, 256)
    vid = [bt.add_volume(v) for v in vols]
    assert vid == [0, 1]
    for r in reqs:
        bt.submit(r["id"], r["volume"], r["desc"], r["lut"])
    done = {}
    while len(done) < len(reqs):
        bt.progress(True)
        for c in bt.poll():
            done[c["id"]] = c
    st = bt.stats()
    print("batcher stats", st)
    assert st["completed"] == len(reqs) and st["rejected"] == 0 and st["batches"] >= len(reqs) // 4
    for r in reqs:
        W, H = r["desc"].state.winwidth, r["desc"].state.winheight
        tf = ctx.tf(r["lut"])
        fr = ctx.frame(W, H)
        ctx.render(vols[r["volume"]], tf, r["desc"], fr)
        alone = fr.download()
        s_alone = fr.stats()["samples"]
        fr.close()
        tf.close()
        c = done[r["id"]]
        assert c["status"] == 0
        assert np.array_equal(c["image"], alone), r["id"]
        assert c["samples"] == s_alone
    bt.close()


def test_budget_admission_and_rejection(ctx, vols):
    vbytes = sum(v.info()["bytes"] for v in vols)
    one = 256 * 256 * 4 + 4096 + 1024  # frame + LUT + arguments of one 256^2 request
    # room for one request at a time (plus the volumes' macrocell bytes)
    bt = ctx.batcher(vbytes + 20000 + int(1.5 * one), 4, 512, 512)
    for v in vols:
        bt.add_volume(v)
    reqs = serving.make_requests(6, 2, seed=12, tf_seed=2005, resolutions=[(256, 256)], skip=False)
    for r in reqs:
        bt.submit(r["id"], r["volume"], r["desc"], r["lut"])
    st = ivr.orbit_state(1, 512, 512)  # 1 MiB frame: more than the budget ever frees
    bt.submit(99, 0, ivr.config_desc(st), configs.tf_lut(2))
    done = {}
    max_admitted_live = 0
    while len(done) < 7:
        bt.progress(False)
        s = bt.stats()
        max_admitted_live = max(max_admitted_live, s["admitted"] - s["completed"])
        for c in bt.poll():
            done[c["id"]] = c
        bt.progress(True)
        for c in bt.poll():
            done[c["id"]] = c
    assert done[99]["status"] == -5  # IVR_ERR_BUDGET
    assert all(done[r["id"]]["status"] == 0 for r in reqs)
    assert max_admitted_live <= 1
    bt.close()


def test_poisson_driver_reports_latency(ctx, vols):
    bt = ctx.batcher(1 << 34, 8, 256, 256)
    for v in vols:
        bt.add_volume(v)
    reqs = serving.make_requests(40, 2, seed=13, tf_seed=2005, resolutions=[(128, 128), (256, 256)])
    res = serving.run_poisson(bt, reqs, rate=200.0, seed=4000)
    print({k: v for k, v in res.items() if k != "records"})
    assert res["ok"] == 40 and res["p99_ms"] >= res["p50_ms"] > 0
    bt.close()


def test_surrogate_requests_in_the_mix(ctx, vols):
    """requests served from a trained surrogate (with and without fallback
    bricks, skipping on and off) next to plain requests: each image and its
    sample count equal ivr_surrogate_render of the same request alone"""
    vol = vols[0]
    sg = ctx.surrogate(vol, log2_table=12)
    sg.train(vol, 20, 1 << 16)
    flags = np.zeros(sg.nbricks, np.uint8)
    flags[::3] = 1  # every third brick renders from the plain volume
    bt = ctx.batcher(1 << 34, 4, 256, 256)
    assert bt.add_volume(vol) == 0
    sg.set_fallback(flags)
    with pytest.raises(ivr.IvrError):
        bt.add_surrogate(sg)  # fallback bricks need the plain volume
    assert bt.add_surrogate(sg, vol) == 1
    sg.set_fallback(np.zeros(sg.nbricks, np.uint8))
    assert bt.add_surrogate(sg) == 2  # no fallback: the network alone
    sizes = [(64, 48), (128, 96), (96, 96)]
    reqs = serving.make_requests(12, 1, seed=21, tf_seed=2005, resolutions=sizes)
    for k, r in enumerate(reqs):
        r["volume"] = k % 3
        r["desc"].skip = k % 2 == 1
        bt.submit(r["id"], r["volume"], r["desc"], r["lut"])
    done = {}
    while len(done) < len(reqs):
        bt.progress(True)
        for c in bt.poll():
            done[c["id"]] = c
    for r in reqs:
        W, H = r["desc"].state.winwidth, r["desc"].state.winheight
        tf = ctx.tf(r["lut"])
        fr = ctx.frame(W, H)
        if r["volume"] == 0:
            ctx.render(vol, tf, r["desc"], fr)
        else:
            sg.set_fallback(flags if r["volume"] == 1 else np.zeros(sg.nbricks, np.uint8))
            sg.render(vol if r["volume"] == 1 else None, tf, r["desc"], fr)
        alone = fr.download()
        s_alone = fr.stats()["samples"]
        fr.close()
        tf.close()
        c = done[r["id"]]
        assert c["status"] == 0
        assert np.array_equal(c["image"], alone), (r["id"], r["volume"])
        assert c["samples"] == s_alone
    bt.close()
    sg.close()
