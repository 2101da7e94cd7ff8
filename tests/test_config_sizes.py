"""The learned-surrogate and steering configurations at their own sizes on
the GPU: the surrogate of the 512^3 fp32 volume (8 bricks x 16 levels x 2^19
x 2 features, MLP 32-64-64-1 on bf16 MFMA) trained by the project's kernels
and rendered at 1920x1080, and the steering service with the 256^3 fp32,
512^3 fp32 and 1024^3 fp16 volumes resident and requests at every
resolution of the mix."""
import os
import tempfile

import numpy as np
import pytest

import ivr_amd as ivr
import oracle
from ivr_amd import configs, serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2_volume(ctx):
    cfg = configs.WORKLOADS[2]
    vol = ctx.volume(*cfg["dims"], cfg["dtype"], brick=cfg["brick"])
    vol.generate(configs.gen_params(2))
    yield vol
    vol.close()


def test_surrogate_of_the_512_volume_at_1080p(ctx, c2_volume):
    """80 Adam steps of 2^24 samples; the 1920x1080 surrogate frame is at
    least 40 dB from the plain GPU render; the per-brick flags are exactly
    the bricks below 40 dB; a frame with every brick falling back is the
    plain frame bit for bit; the renderer's network value at points of two
    bricks matches the oracle on the trained, fp16-rounded tables; the same
    network in a context without helper lanes (IVR_SG_HELPERS=0) renders
    the same frame and counts"""
    vol = c2_volume
    sg = ctx.surrogate(vol)
    assert sg.nbricks == 8 and sg.T == 1 << 19
    losses = [sg.train(vol, 10, 1 << 24) for _ in range(8)]
    print("losses per 10 steps:", " ".join("%.3e" % v for v in losses), flush=True)
    assert losses[-1] < 0.05 * losses[0]
    W, H = 1920, 1080
    tf = ctx.tf(configs.tf_lut(2))
    desc = ivr.config_desc(ivr.orbit_state(3004, W, H), skip=True)
    fr = ctx.frame(W, H)
    ctx.render(vol, tf, desc, fr)
    plain = fr.download()
    sg.set_fallback(np.zeros(sg.nbricks, np.uint8))
    sg.render(vol, tf, desc, fr)
    img = fr.download()
    p_all = ivr.psnr(img, plain)
    bpsnr, nfall = sg.evaluate(vol, tf, desc, threshold_db=40.0)
    print("PSNR %.2f dB, SSIM %.4f; per-brick PSNR %s; %d bricks fall back"
          % (p_all, ivr.ssim(img, plain), " ".join("%.1f" % v for v in bpsnr), nfall), flush=True)
    assert nfall == int((bpsnr < 40.0).sum())
    sg.render(vol, tf, desc, fr)
    p_fb = ivr.psnr(fr.download(), plain)
    assert p_fb >= 40.0 and p_fb >= p_all - 1e-9
    sg.set_fallback(np.ones(sg.nbricks, np.uint8))
    sg.render(vol, tf, desc, fr)
    assert np.array_equal(fr.download(), plain)
    # the network value at points of two bricks against the oracle
    rng = np.random.default_rng(44)
    n = 256
    xyz = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    bricks = np.array([0, 5], np.int32)[rng.integers(0, 2, n)]
    y = sg.eval_points(xyz, bricks)
    yo = np.zeros(n)
    for b in (0, 5):
        mlp, table = sg.get(brick=b, table=True)
        sel = bricks == b
        yo[sel] = oracle.sg_eval(mlp, table.astype(np.float16).astype(np.float32), sg.T,
                                 xyz[sel].astype(np.float64))
    err = np.abs(y - yo).max()
    print("trained network: max |y_render - y_oracle| = %.3e" % err)
    assert err < 2e-3 * max(1.0, np.abs(yo).max())
    # helper lanes change no frame: the same checkpoint without them
    sg.set_fallback(np.zeros(sg.nbricks, np.uint8))
    sg.render(vol, tf, desc, fr)
    img, st = fr.download(), fr.stats()
    path = os.path.join(tempfile.mkdtemp(), "c4.ckpt")
    sg.save(path)
    os.environ["IVR_SG_HELPERS"] = "0"
    try:
        ctx2 = ivr.Context(0)
    finally:
        del os.environ["IVR_SG_HELPERS"]
    cfg = configs.WORKLOADS[2]
    vol2 = ctx2.volume(*cfg["dims"], cfg["dtype"], brick=cfg["brick"])
    vol2.generate(configs.gen_params(2))
    sg2 = ctx2.surrogate(vol2)
    sg2.load(path)
    sg2.set_fallback(np.zeros(sg2.nbricks, np.uint8))
    tf2 = ctx2.tf(configs.tf_lut(2))
    fr2 = ctx2.frame(W, H)
    sg2.render(vol2, tf2, desc, fr2)
    st2 = fr2.stats()
    print("lane utilisation with helpers %.3f, without %.3f"
          % (st["visited"] / st["evaluations"], st2["visited"] / st2["evaluations"]))
    assert np.array_equal(fr2.download(), img)
    assert (st["samples"], st["visited"]) == (st2["samples"], st2["visited"])
    for o in (fr2, tf2, sg2, vol2):
        o.close()
    ctx2.close()
    fr.close()
    tf.close()
    sg.close()


def test_steering_mix_at_config_sizes(ctx, c2_volume):
    """256^3 fp32, 512^3 fp32 and 1024^3 fp16 resident; requests at 256^2,
    512^2, 1024^2 and 1920x1080 over all three volumes with jittered TFs and
    skipping on and off, behind an HBM budget of about two and a half of the
    largest frames beyond the residents, so requests queue for memory; every
    image and sample count equals the same request rendered alone"""
    v256 = ctx.volume(256, 256, 256, ivr.IVR_F32, brick=256)
    v256.generate(configs.gen_params(2))
    cfg3 = configs.WORKLOADS[3]
    v1024 = ctx.volume(*cfg3["dims"], cfg3["dtype"], brick=cfg3["brick"])
    v1024.generate(configs.gen_params(3))
    vols = [v256, c2_volume, v1024]
    probe = ctx.batcher(1 << 40, 1, 16, 16)
    for v in vols:
        probe.add_volume(v)
    base = probe.stats()["bytes_in_use"]
    probe.close()
    big = 1920 * 1080 * 4 + 300000  # largest frame + LUT + arguments + macrocell codes
    budget = base + int(2.5 * big)
    bt = ctx.batcher(budget, 16, 1920, 1080)
    assert [bt.add_volume(v) for v in vols] == [0, 1, 2]
    reqs = serving.make_requests(24, 3, seed=4200, tf_seed=2005)
    for k, r in enumerate(reqs):  # every (volume, resolution) pair, skipping on and off
        w, h = serving.RESOLUTIONS[k % 4]
        r["volume"] = (k // 4) % 3
        r["desc"] = ivr.config_desc(ivr.orbit_state(5000 + k, w, h), skip=k % 2 == 0)
    for r in reqs:
        bt.submit(r["id"], r["volume"], r["desc"], r["lut"])
    done, max_live, max_pending = {}, 0, 0
    while len(done) < len(reqs):
        bt.progress(False)
        s = bt.stats()
        assert s["bytes_in_use"] <= budget
        max_live = max(max_live, s["admitted"] - s["completed"] - s["rejected"])
        max_pending = max(max_pending, s["pending"])
        bt.progress(True)
        for c in bt.poll():
            done[c["id"]] = c
    print("admitted at once at most %d of %d requests; %d waited for memory at most" %
          (max_live, len(reqs), max_pending), flush=True)
    assert max_pending > 0 and max_live < len(reqs)
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
        assert c["status"] == 0, r["id"]
        assert np.array_equal(c["image"], alone), (r["id"], r["volume"], W, H)
        assert c["samples"] == s_alone
    bt.close()
    v256.close()
    v1024.close()
