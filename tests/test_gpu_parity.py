"""GPU parity of the HIP render path (libivr.so through its C-ABI) against
the reference renderer's golden vectors and the CPU oracle.

Bar: compat mode bit-exact with the compiled reference's RGBA32F images;
config mode within 1/255 per RGBA8 channel of the fp32 oracle (it is
expected to be bit-exact: both compute in the same float order) with RMSE
and the FP64 oracle error reported; invariances (bricking, macrocell
skipping, determinism) bit-exact.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import ivr_amd as ivr
import oracle
from ivr_amd import configs

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(G, "golden_meta.json")))


def _render(ctx, vol, lut, desc, W, H):
    tf = ctx.tf(lut)
    fr = ctx.frame(W, H, keep_rgba32=True)
    ctx.render(vol, tf, desc, fr)
    img8, img32 = fr.download(rgba32=True)
    st = fr.stats()
    fr.close()
    tf.close()
    return img8, img32, st


def _oracle_sort_last(host, st, lut, B, rows=None):
    """oracle tiles of every brick (near to far by the sum of axis distances
    to the eye's brick, ties by id) composited by compositetiles"""
    import ctypes as C
    dims = (host.shape[2], host.shape[1], host.shape[0])
    v = oracle.view_of(st)
    cam = np.zeros(23, np.float32)
    rect = (C.c_int * 4)()
    bmin, bmax = oracle.box(*dims)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    oracle.lib.ivro_camera(fp(bmin), fp(bmax), C.byref(v), fp(cam), rect)
    eb = [int(np.floor((cam[r] - bmin[r]) * (dims[r] / (bmax[r] - bmin[r])) / B)) for r in range(3)]
    nb = [(dims[r] - 2) // B + 1 for r in range(3)]
    ids = sorted([(x, y, z) for z in range(nb[2]) for y in range(nb[1]) for x in range(nb[0])],
                 key=lambda b: (sum(abs(b[r] - eb[r]) for r in range(3)), b[0] + nb[0] * (b[1] + nb[1] * b[2])))
    tiles, s_ref = [], 0
    for b in ids:
        t, s = oracle.render(host, v, lut, oracle.params(brick=B, brick_id=b), rows=rows)
        tiles.append(t)
        s_ref += s
    return oracle.composite_tiles(np.stack(tiles)), s_ref


def _report(tag, img8, ref32, img32=None):
    ref8 = oracle.quantize(ref32)
    d8 = np.abs(img8.astype(int) - ref8.astype(int))
    rmse = float(np.sqrt(np.mean((img8.astype(float) - ref8.astype(float)) ** 2)))
    same = None if img32 is None else float(np.mean(img32.view(np.uint32) == ref32.view(np.uint32)))
    print("%s: max|d8|=%d rmse8=%.4f bitwise-identical=%s" % (tag, d8.max(), rmse, same), flush=True)
    return d8.max(), rmse, same


@pytest.fixture(scope="module")
def c1_host():
    gp = configs.gen_params(1)
    return ivr.generate_host(gp, 64, 64, 64)


def test_compat_matches_compiled_reference(ctx):
    """GPU compat mode == the reference's createlocalimage, bit for bit."""
    gp = configs.gen_params(1)
    gp.value_scale = 255.0
    host = ivr.generate_host(gp, 64, 64, 64)
    assert hashlib.sha256(host.tobytes()).hexdigest() == META["sha256"]
    vol = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=64)
    vol.upload(host)
    z = np.load(os.path.join(G, "compat_images.npz"))
    luts = {"tf5": z["lut_tf5"], "medical": z["lut_medical"], "jet": z["lut_jet"]}
    W, H = META["width"], META["height"]
    for k, case in enumerate(META["compat_cases"]):
        print("compat case", k, case, flush=True)
        st = ivr.state_from_angles(case["az"], case["el"], W, H)
        st.alpha = case["alpha"]
        d = ivr.compat_desc(st, xfer=case["xfer"], maxlutindx=case["maxlutindx"])
        img8, img32, stats = _render(ctx, vol, luts[case["lut"]], d, W, H)
        ref = z["images"][k]
        assert np.array_equal(img32.view(np.uint32), ref.view(np.uint32)), case
        assert np.array_equal(img8, oracle.quantize(ref))
    vol.close()


def _state_of(row, W, H):
    """ivr_amd.State from a stored row: mvm[16], pm[16], vv[4], newconst[6], alpha"""
    st = ivr.State()
    st.winwidth, st.winheight = W, H
    st.mvm[:] = [float(v) for v in row[0:16]]
    st.pm[:] = [float(v) for v in row[16:32]]
    st.vv[:] = [float(v) for v in row[32:36]]
    st.newconst[:] = [float(v) for v in row[36:42]]
    st.alpha = float(row[42])
    return st


def test_cut_planes_pan_and_wide_rect(ctx):
    """STATE cut planes, a panned eye and a box projected past the viewport:
    compat mode == the compiled reference's frames bit for bit; config mode
    == the oracle bit for bit, fused and sort-last, skipping on and off, with
    2^3 bricks"""
    gp = configs.gen_params(1)
    host = ivr.generate_host(gp, 64, 64, 64)
    gp.value_scale = 255.0
    host255 = ivr.generate_host(gp, 64, 64, 64)
    z = np.load(os.path.join(G, "compat_cuts.npz"))
    W, H = META["width"], META["height"]
    vol255 = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=64)
    vol255.upload(host255)
    vol = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=32)
    vol.upload(host)
    lut = np.ascontiguousarray(z["lut"])
    for k, row in enumerate(z["states"]):
        print("cut/pan case", k, flush=True)
        st = _state_of(row, W, H)
        img8, img32, _ = _render(ctx, vol255, lut, ivr.compat_desc(st, xfer=3, maxlutindx=255), W, H)
        assert np.array_equal(img32.view(np.uint32), z["images"][k].view(np.uint32)), k
        ref32, s_ref = oracle.render(host, oracle.view_of(st), lut, oracle.params())
        sl32, s_sl = _oracle_sort_last(host, st, lut, 32)
        for skip in (False, True):
            f8, f32, fs = _render(ctx, vol, lut, ivr.config_desc(st, skip=skip), W, H)
            assert np.array_equal(f32.view(np.uint32), ref32.view(np.uint32)) and fs["samples"] == s_ref, (k, skip)
            l8, l32, ls = _render(ctx, vol, lut, ivr.config_desc(st, skip=skip, composite=ivr.SORT_LAST), W, H)
            assert np.array_equal(l32.view(np.uint32), sl32.view(np.uint32)) and ls["samples"] == s_sl, (k, skip)
    vol.close()
    vol255.close()


def test_generator_gpu_equals_host(ctx):
    for (kind_cfg, n, brick) in ((1, 64, 64), (2, 96, 32)):
        gp = configs.gen_params(kind_cfg)
        host = ivr.generate_host(gp, n, n, n)
        vol = ctx.volume(n, n, n, ivr.IVR_F32, brick=brick)
        vol.generate(gp)
        dev = vol.download()
        vol.close()
        assert np.array_equal(dev.view(np.uint32), host.view(np.uint32)), kind_cfg
        if kind_cfg == 2:
            zf = float(np.mean(host == 0.0))
            assert abs(zf - 0.4) < 0.01, zf


def test_config_c1_matches_oracle(ctx, c1_host):
    vol = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=64)
    vol.upload(c1_host)
    lut = configs.tf_lut(1)
    st = configs.state(1)
    d = ivr.config_desc(st)
    img8, img32, stats = _render(ctx, vol, lut, d, 256, 256)
    ref32, s_ref = oracle.render(c1_host, oracle.view_of(st), lut, oracle.params())
    ref64, _ = oracle.render(c1_host, oracle.view_of(st), lut, oracle.params(), f64=True)
    m, rmse, same = _report("C1 config vs fp32 oracle", img8, ref32, img32)
    e64 = float(np.abs(img32.astype(np.float64) - ref64).max())
    print("C1 max |gpu - fp64 oracle| = %.3e (RGBA32F)" % e64, flush=True)
    assert m <= 1
    assert stats["samples"] == s_ref
    assert same == 1.0
    # the download into a caller's reused buffer is the same image
    tf = ctx.tf(lut)
    fr = ctx.frame(256, 256)
    ctx.render(vol, tf, d, fr)
    buf = np.full((256, 256, 4), 7, np.uint8)
    assert fr.download(out=buf) is buf
    assert np.array_equal(buf, img8)
    with pytest.raises(ValueError):
        fr.download(out=np.empty((256, 255, 4), np.uint8))
    fr.close()
    tf.close()
    vol.close()


def test_config_fp16_matches_oracle(ctx, c1_host):
    vol = ctx.volume(64, 64, 64, ivr.IVR_F16, brick=32)
    vol.upload(c1_host)
    h16 = c1_host.astype(np.float16).astype(np.float32)
    assert np.array_equal(vol.download(), h16)
    lut = configs.tf_lut(1)
    st = configs.state(1, 192, 160)
    img8, img32, stats = _render(ctx, vol, lut, ivr.config_desc(st), 192, 160)
    ref32, s_ref = oracle.render(h16, oracle.view_of(st), lut, oracle.params())
    m, rmse, same = _report("fp16 config vs oracle", img8, ref32, img32)
    assert m <= 1 and stats["samples"] == s_ref
    vol.close()


def test_bricking_and_skipping_are_invisible(ctx):
    """2x2x2 / 4x4x4 bricks == one brick; macrocell skip on == off, bit for bit"""
    gp = configs.gen_params(2)
    n = 128
    host = ivr.generate_host(gp, n, n, n)
    lut = configs.tf_lut(2)
    st = configs.state(2, 200, 176)
    outs = []
    for brick, skip in ((128, False), (64, False), (32, True), (128, True)):
        vol = ctx.volume(n, n, n, ivr.IVR_F32, brick=brick)
        vol.upload(host)
        outs.append(_render(ctx, vol, lut, ivr.config_desc(st, skip=skip), 200, 176))
        vol.close()
    base8, base32, base_st = outs[0]
    for img8, img32, s in outs[1:]:
        assert np.array_equal(img32.view(np.uint32), base32.view(np.uint32))
        assert s["samples"] == base_st["samples"]
    print("samples %d, fetched with skipping %d" % (outs[3][2]["samples"], outs[3][2]["visited"]))
    assert outs[3][2]["visited"] < outs[3][2]["samples"]
    ref32, s_ref = oracle.render(host, oracle.view_of(st), lut, oracle.params())
    m, rmse, same = _report("128^3 noise config vs oracle", base8, ref32, base32)
    assert same == 1.0 and m == 0 and s_ref == base_st["samples"]


@pytest.mark.parametrize("composite,B", [("SORT_LAST", 32), ("SORT_LAST", 16), ("SORT_LAST_TILES", 16)])
def test_sort_last_matches_oracle_tiles(ctx, composite, B):
    """per-brick partials composited in visibility order == the oracle's tile
    renders composited by compositetiles, bit for bit: the default per-run
    workgroups (k_runs + k_combine_runs; 4^3 bricks with run counts), the
    one-launch walks of a context made with IVR_SL_SPLIT=0 (2^3 bricks:
    partials in LDS; 4^3 bricks: runs walked from the far end), and the
    materialised tiles; macrocell skipping invisible"""
    gp = configs.gen_params(2)
    n = 64
    host = ivr.generate_host(gp, n, n, n)
    lut = configs.tf_lut(2)
    st = configs.state(2, 96, 80)
    vol = ctx.volume(n, n, n, ivr.IVR_F32, brick=B)
    vol.upload(host)
    mode = getattr(ivr, composite)
    img8, img32, stats = _render(ctx, vol, lut, ivr.config_desc(st, composite=mode), 96, 80)
    if composite == "SORT_LAST":
        s8, s32, sstats = _render(ctx, vol, lut, ivr.config_desc(st, composite=mode, skip=True), 96, 80)
        assert np.array_equal(s32.view(np.uint32), img32.view(np.uint32))
        assert sstats["samples"] == stats["samples"] and sstats["visited"] <= stats["visited"]
        print("sort-last with skipping: %d of %d samples fetched" % (sstats["visited"], stats["visited"]))
        # the single-launch walk of all runs per ray (IVR_SL_SPLIT=0) gives the same image
        os.environ["IVR_SL_SPLIT"] = "0"
        try:
            wctx = ivr.Context(0)
        finally:
            del os.environ["IVR_SL_SPLIT"]
        wvol = wctx.volume(n, n, n, ivr.IVR_F32, brick=B)
        wvol.upload(host)
        for skip in (False, True):
            w8, w32, wstats = _render(wctx, wvol, lut, ivr.config_desc(st, composite=mode, skip=skip), 96, 80)
            assert np.array_equal(w32.view(np.uint32), img32.view(np.uint32)) and wstats["samples"] == stats["samples"]
        wvol.close()
        wctx.close()
    vol.close()
    ref32, s_ref = _oracle_sort_last(host, st, lut, B)
    m, rmse, same = _report("sort-last vs oracle tiles", img8, ref32, img32)
    assert same == 1.0 and stats["samples"] == s_ref
    # and within the ERT bound of the fused (global ERT) render
    vol = ctx.volume(n, n, n, ivr.IVR_F32, brick=B)
    vol.upload(host)
    f8, f32, _ = _render(ctx, vol, lut, ivr.config_desc(st), 96, 80)
    vol.close()
    assert np.abs(f32 - img32).max() <= 0.0101


def _sparse_volume(nx, ny, nz, seed):
    """zeros with a few gaussian blobs: large transparent regions, so the
    skip jumps over nodes of up to 8^3 macrocells"""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vol = np.zeros((nz, ny, nx), np.float32)
    for _ in range(6):
        c = rng.uniform(0.15, 0.85, 3) * (nx, ny, nz)
        r2 = ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / rng.uniform(200.0, 600.0)
        vol += np.exp(-r2).astype(np.float32)
    return np.minimum(vol, 1.0).astype(np.float32)


def _sparse_lut():
    lut = np.zeros((256, 4), np.float32)
    t = np.linspace(0.0, 1.0, 256, dtype=np.float32)
    lut[:, 0], lut[:, 1], lut[:, 2] = t, 1.0 - t, 0.5 + 0.5 * t
    lut[100:, 3] = np.linspace(0.05, 0.6, 156, dtype=np.float32)
    return lut


@pytest.mark.parametrize("B", [32, 64])
def test_hierarchical_skip_on_sparse_volume(ctx, B):
    """a mostly transparent 160x144x176 volume: skipping over empty nodes of
    1..8 macrocells per axis (crossing brick faces) is invisible, fused and
    sort-last, bit for bit against the oracle, from several views"""
    nx, ny, nz = 160, 144, 176
    host = _sparse_volume(nx, ny, nz, 7)
    lut = _sparse_lut()
    W, H = 96, 88
    vol = ctx.volume(nx, ny, nz, ivr.IVR_F32, brick=B)
    vol.upload(host)
    for seed in (11, 12, 13):
        st = configs.state(2, W, H, seed=seed)
        ref32, s_ref = oracle.render(host, oracle.view_of(st), lut, oracle.params())
        f8, f32, fs = _render(ctx, vol, lut, ivr.config_desc(st, skip=True), W, H)
        _report("sparse fused skip B=%d view %d" % (B, seed), f8, ref32, f32)
        assert np.array_equal(f32.view(np.uint32), ref32.view(np.uint32)) and fs["samples"] == s_ref
        assert fs["visited"] < fs["samples"] // 4
        sl32, s_sl = _oracle_sort_last(host, st, lut, B)
        l8, l32, ls = _render(ctx, vol, lut, ivr.config_desc(st, skip=True, composite=ivr.SORT_LAST), W, H)
        _report("sparse sort-last skip B=%d view %d" % (B, seed), l8, sl32, l32)
        assert np.array_equal(l32.view(np.uint32), sl32.view(np.uint32)) and ls["samples"] == s_sl
    vol.close()


@pytest.mark.parametrize("composite", ["SORT_LAST", "SORT_LAST_TILES"])
def test_fp16_sort_last_with_skipping_matches_oracle(ctx, composite):
    """an fp16 volume of 4x3x5 bricks (the C3 storage, at test size), sort-last
    with macrocell skipping on and off, bit for bit against the oracle's
    per-brick tiles of the fp16-rounded field"""
    gp = configs.gen_params(3)
    nx, ny, nz = 96, 80, 72
    host = ivr.generate_host(gp, nx, ny, nz)
    h16 = host.astype(np.float16).astype(np.float32)
    lut = configs.tf_lut(3)
    W, H = 112, 96
    st = configs.state(3, W, H)
    vol = ctx.volume(nx, ny, nz, ivr.IVR_F16, brick=16 if composite == "SORT_LAST_TILES" else 32)
    vol.upload(host)
    B = vol.info()["brick"]
    ref32, s_ref = _oracle_sort_last(h16, st, lut, B)
    mode = getattr(ivr, composite)
    for skip in (False, True):
        img8, img32, s = _render(ctx, vol, lut, ivr.config_desc(st, composite=mode, skip=skip), W, H)
        _report("fp16 %s B=%d skip=%d" % (composite, B, skip), img8, ref32, img32)
        assert np.array_equal(img32.view(np.uint32), ref32.view(np.uint32)) and s["samples"] == s_ref
    vol.close()


@pytest.mark.parametrize("B", [16, 32])
def test_ragged_sizes_match_oracle(ctx, B):
    """a 53x37x29 volume (no axis a multiple of the brick; bricks partly
    filled, clamped tails) into a 101x67 frame (partial pixel blocks), fused
    and sort-last, skipping on and off, bit for bit against the oracle"""
    gp = configs.gen_params(2)
    nx, ny, nz = 53, 37, 29
    host = ivr.generate_host(gp, nx, ny, nz)
    lut = configs.tf_lut(2)
    W, H = 101, 67
    st = configs.state(2, W, H)
    vol = ctx.volume(nx, ny, nz, ivr.IVR_F32, brick=B)
    vol.upload(host)
    assert vol.info()["nbricks"] == tuple((d - 2) // B + 1 for d in (nx, ny, nz))
    ref32, s_ref = oracle.render(host, oracle.view_of(st), lut, oracle.params())
    f8, f32, fs = _render(ctx, vol, lut, ivr.config_desc(st), W, H)
    _report("ragged fused B=%d" % B, f8, ref32, f32)
    assert np.array_equal(f32.view(np.uint32), ref32.view(np.uint32)) and fs["samples"] == s_ref
    k8, k32, ks = _render(ctx, vol, lut, ivr.config_desc(st, skip=True), W, H)
    assert np.array_equal(k32.view(np.uint32), f32.view(np.uint32)) and ks["samples"] == s_ref
    sl32, s_sl = _oracle_sort_last(host, st, lut, B)
    for skip in (False, True):
        l8, l32, ls = _render(ctx, vol, lut, ivr.config_desc(st, skip=skip, composite=ivr.SORT_LAST), W, H)
        _report("ragged sort-last B=%d skip=%d" % (B, skip), l8, sl32, l32)
        assert np.array_equal(l32.view(np.uint32), sl32.view(np.uint32)) and ls["samples"] == s_sl
    vol.close()


def test_minimum_sizes(ctx):
    """the smallest volume (2^3) and frames of one pixel and one row"""
    gp = configs.gen_params(1)
    host = ivr.generate_host(gp, 2, 2, 2)
    lut = configs.tf_lut(1)
    vol = ctx.volume(2, 2, 2, ivr.IVR_F32, brick=16)
    vol.upload(host)
    for (W, H) in ((1, 1), (33, 1), (1, 17)):
        st = configs.state(1, W, H)
        for comp in (ivr.FUSED, ivr.SORT_LAST):
            img8, img32, s = _render(ctx, vol, lut, ivr.config_desc(st, composite=comp), W, H)
            ref32, s_ref = oracle.render(host, oracle.view_of(st), lut, oracle.params())
            assert np.array_equal(img32.view(np.uint32), ref32.view(np.uint32)) and s["samples"] == s_ref, (W, H)
    vol.close()


def test_c2_full_size_rows_match_oracle(ctx):
    """C2 at its full size: 512^3 generated on the GPU, 1024^2 frame; a band
    of rows checked against the oracle (sort-last: per-brick tiles), every 8th
    row of the fused and the sort-last frame against the one-pass oracle,
    determinism of repeated renders"""
    cfg = configs.WORKLOADS[2]
    vol = ctx.volume(*cfg["dims"], cfg["dtype"], brick=cfg["brick"])
    vol.generate(configs.gen_params(2))
    info = vol.info()
    assert info["nbricks"] == (2, 2, 2)
    host = vol.download()
    assert abs(float(np.mean(host == 0.0)) - 0.4) < 0.005
    lut = configs.tf_lut(2)
    st = configs.state(2)
    d = ivr.config_desc(st)
    a8, a32, s1 = _render(ctx, vol, lut, d, 1024, 1024)
    b8, b32, s2 = _render(ctx, vol, lut, d, 1024, 1024)
    assert np.array_equal(a32.view(np.uint32), b32.view(np.uint32)) and s1["samples"] == s2["samples"]
    rows = (448, 512)
    ref32, _ = oracle.render(host, oracle.view_of(st), lut, oracle.params(), rows=rows)
    ref64, _ = oracle.render(host, oracle.view_of(st), lut, oracle.params(), rows=rows, f64=True)
    sl = slice(*rows)
    m, rmse, same = _report("C2 rows %s vs oracle" % (rows,), a8[sl], ref32[sl], a32[sl])
    print("C2 max |gpu - fp64 oracle| = %.3e" % float(np.abs(a32[sl] - ref64[sl]).max()), flush=True)
    assert same == 1.0
    assert (a8[..., 3] > 0).mean() > 0.2
    # the C2 headline composite: sort-last over the 8 bricks
    ds = ivr.config_desc(st, composite=ivr.SORT_LAST)
    c8, c32, s3 = _render(ctx, vol, lut, ds, 1024, 1024)
    rows = (480, 512)
    sl = slice(*rows)
    ref32, _ = _oracle_sort_last(host, st, lut, cfg["brick"], rows=rows)
    m, rmse, same = _report("C2 sort-last rows %s vs oracle tiles" % (rows,), c8[sl], ref32[sl], c32[sl])
    assert same == 1.0
    # every 8th row of both frames against the one-pass oracle
    spread = list(range(0, 1024, 8))
    for tag, img8, img32, p in (("fused", a8, a32, oracle.params()),
                                ("sort-last", c8, c32, oracle.params(brick=cfg["brick"], sort_last=True))):
        ref = np.stack([oracle.render(host, oracle.view_of(st), lut, p, rows=(r, r + 1))[0][r] for r in spread])
        m, rmse, same = _report("C2 %s, %d rows spread over the frame" % (tag, len(spread)), img8[spread], ref,
                                img32[spread])
        assert same == 1.0
    print("C2 samples per frame: fused %d, sort-last %d" % (s1["samples"], s3["samples"]), flush=True)
    vol.close()


def test_c3_full_size_rows_match_oracle(ctx):
    """C3 at its full size: 1024^3 fp16 generated on the GPU (4x4x4 bricks,
    70 % zero), 1920x1080 with macrocell skipping; a band of rows of the
    sort-last and the fused frame bit-exact against the oracle (per-brick
    tiles composited, and one pass), and every 5th row of the frame against
    the one-pass oracle"""
    cfg = configs.WORKLOADS[3]
    vol = ctx.volume(*cfg["dims"], cfg["dtype"], brick=cfg["brick"])
    vol.generate(configs.gen_params(3))
    assert vol.info()["nbricks"] == (4, 4, 4)
    host = vol.download()
    assert abs(float(np.mean(host == 0.0)) - 0.7) < 0.005
    lut = configs.tf_lut(3)
    st = configs.state(3)
    W, H = cfg["image"]
    rows = (532, 548)
    sl = slice(*rows)
    for comp in (ivr.SORT_LAST, ivr.FUSED):
        print("C3 full size, composite", comp, flush=True)
        img8, img32, _ = _render(ctx, vol, lut, ivr.config_desc(st, skip=True, composite=comp), W, H)
        if comp == ivr.SORT_LAST:
            ref32, _ = _oracle_sort_last(host, st, lut, cfg["brick"], rows=rows)
        else:
            ref32, _ = oracle.render(host, oracle.view_of(st), lut, oracle.params(), rows=rows)
        m, rmse, same = _report("C3 rows %s composite %d" % (rows, comp), img8[sl], ref32[sl], img32[sl])
        assert same == 1.0
        assert (img8[sl][..., 3] > 0).mean() > 0.05
        # every 5th row of the whole frame against the one-pass oracle (sort-last: each
        # ray's per-brick runs combined far to near, equal to the tiles by the oracle's own test)
        p = oracle.params(brick=cfg["brick"], sort_last=True) if comp == ivr.SORT_LAST else oracle.params()
        spread = list(range(0, H, 5))
        ref = np.stack([oracle.render(host, oracle.view_of(st), lut, p, rows=(r, r + 1))[0][r] for r in spread])
        m, rmse, same = _report("C3 %d rows spread over the frame, composite %d" % (len(spread), comp),
                                img8[spread], ref, img32[spread])
        assert same == 1.0
    vol.close()


def test_volume_over_4gib_matches_oracle(ctx):
    """a 1280^3 fp32 volume (125 bricks of 256^3, 8.6 GB of bricks: past the
    4 GiB raw-buffer range, so the kernels take their 64-bit addressing path)
    into a 512x512 frame with skipping; a band of rows of the fused and the
    sort-last frame bit-exact against the oracle"""
    n = 1280
    vol = ctx.volume(n, n, n, ivr.IVR_F32, brick=256)
    vol.generate(configs.gen_params(3))
    info = vol.info()
    assert info["nbricks"] == (5, 5, 5) and info["bytes"] > (4 << 30)
    host = vol.download()
    lut = configs.tf_lut(3)
    st = configs.state(3, 512, 512)
    rows = (248, 264)
    sl = slice(*rows)
    for comp in (ivr.FUSED, ivr.SORT_LAST):
        print("over 4 GiB, composite", comp, flush=True)
        img8, img32, _ = _render(ctx, vol, lut, ivr.config_desc(st, skip=True, composite=comp), 512, 512)
        if comp == ivr.SORT_LAST:
            ref32, _ = _oracle_sort_last(host, st, lut, 256, rows=rows)
        else:
            ref32, _ = oracle.render(host, oracle.view_of(st), lut, oracle.params(), rows=rows)
        m, rmse, same = _report("1280^3 rows %s composite %d" % (rows, comp), img8[sl], ref32[sl], img32[sl])
        assert same == 1.0
        assert (img8[sl][..., 3] > 0).mean() > 0.05
    vol.close()


def test_error_paths(ctx, c1_host):
    vol = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=64)
    vol.upload(c1_host)
    tf = ctx.tf(configs.tf_lut(1))
    fr = ctx.frame(64, 64)
    st = configs.state(1, 128, 128)  # state size does not match the frame
    with pytest.raises(ivr.IvrError, match="does not match"):
        ctx.render(vol, tf, ivr.config_desc(st), fr)
    with pytest.raises(ivr.IvrError):
        ctx.volume(64, 64, 64, ivr.IVR_F32, brick=48)
    with pytest.raises(ivr.IvrError):
        ctx.tf(np.zeros((5000, 4), np.float32))
    fr.close()
    tf.close()
    vol.close()
    # allocations that cannot succeed (TiB-sized) fail with IVR_ERR_NOMEM and release what the
    # call had already allocated
    before = ctx.device_bytes()
    with pytest.raises(ivr.IvrError, match="hipMalloc"):
        ctx.batcher(1 << 40, 2, 1 << 20, 1 << 20)
    assert ctx.device_bytes() == before
    with pytest.raises(ivr.IvrError, match="hipMalloc"):
        ctx.volume(8192, 8192, 8192, ivr.IVR_F32, brick=512)
    assert ctx.device_bytes() == before
    small = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=64)  # the context still works
    small.upload(c1_host)
    small.close()
    assert ctx.device_bytes() == before


@pytest.mark.parametrize("dims,f16,brick", [((37, 50, 29), False, 16), ((17, 17, 17), False, 16),
                                            ((40, 33, 70), True, 16), ((128, 96, 160), False, 64)])
def test_gradient_field_matches_oracle(ctx, dims, f16, brick):
    """ivr_volume_gradient (central differences over the bricks, one-sided
    on the faces) == the oracle's gradient3d restatement, bit for bit; ragged
    dims, a last brick owning its tail voxel, fp16 storage"""
    nx, ny, nz = dims
    print("gradient", dims, "f16" if f16 else "f32", "brick", brick, flush=True)
    host = np.random.default_rng(nx).uniform(-1.0, 1.0, (nz, ny, nx)).astype(np.float32)
    vol = ctx.volume(nx, ny, nz, ivr.IVR_F16 if f16 else ivr.IVR_F32, brick=brick)
    vol.upload(host)
    stored = vol.download()  # the values as stored (fp16-rounded for f16)
    got = vol.gradient()
    ref = oracle.gradient(stored)
    vol.close()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_gradient_field_into_device_memory(ctx):
    """the stream-ordered device-destination form of ivr_volume_gradient
    writes the same field as the host form"""
    import ctypes as C
    nx, ny, nz = 70, 41, 33
    host = np.random.default_rng(3).uniform(-1.0, 1.0, (nz, ny, nx)).astype(np.float32)
    vol = ctx.volume(nx, ny, nz, ivr.IVR_F32, brick=32)
    vol.upload(host)
    ref = vol.gradient()
    hip = C.CDLL("libamdhip64.so")
    dst = C.c_void_p()
    nbytes = nx * ny * nz * 16
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    try:
        vol.gradient(device_ptr=dst.value)
        ctx.sync()
        got = np.empty((nz, ny, nx, 4), np.float32)
        assert hip.hipMemcpy(C.c_void_p(got.ctypes.data), dst, C.c_size_t(nbytes), 2) == 0  # to host
    finally:
        hip.hipFree(dst)
    vol.close()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_gradient_field_c2_full_size(ctx):
    """the gradient of the 512^3 generated volume (2x2x2 bricks of 256^3),
    bit-exact with the oracle over every voxel"""
    vol = ctx.volume(512, 512, 512, ivr.IVR_F32, brick=256)
    vol.generate(configs.gen_params(2))
    stored = vol.download()
    got = vol.gradient()
    vol.close()
    ref = oracle.gradient(stored)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
