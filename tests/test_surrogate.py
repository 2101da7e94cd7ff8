"""Learned surrogate: CPU oracle pins (finite differences, level split,
Adam closed form) and, on the GPU, gradient parity of the MFMA training
kernel with the bf16-emulating oracle, training gain, per-brick fallback and
checkpoint round trip."""
import os
import tempfile

import numpy as np
import pytest

import ivr_amd as ivr
import oracle
from ivr_amd import configs

T_SMALL = 1 << 12  # small tables keep the CPU oracle fast


def _params(rng):
    p = np.zeros(oracle.SG_NPARAM, np.float32)
    p[:64 * 32] = rng.uniform(-0.3, 0.3, 64 * 32)
    p[64 * 32:64 * 33] = rng.uniform(-0.1, 0.1, 64)
    o = 64 * 33
    p[o:o + 4096] = rng.uniform(-0.2, 0.2, 4096)
    p[o + 4096:o + 4160] = rng.uniform(-0.1, 0.1, 64)
    p[o + 4160:o + 4224] = rng.uniform(-0.3, 0.3, 64)
    p[-1] = 0.05
    return p


def test_level_resolutions_and_dense_split():
    res = [oracle.lib.ivros_level_res(l, 16, 1.5) for l in range(16)]
    assert res[:6] == [16, 24, 36, 54, 81, 121] and res[-1] == 7006  # floor(16 * 1.5**15) = floor(7006.04)
    T = 1 << 19
    # (N+1)^3 <= T: levels 0-3 dense (17^3 .. 55^3), level 4 (82^3) hashed
    assert oracle.lib.ivros_index(3, 16, 1.5, T, 54, 54, 54) == 54 + 55 * (54 + 55 * 54)
    assert oracle.lib.ivros_index(4, 16, 1.5, T, 81, 81, 81) < T


def test_block_hash_layout():
    """hashed levels: a 4x4x4 block of vertices occupies one 64-entry slot
    chosen by the spatial hash of its block coordinates, at place
    x1 x0 | y0 | z0 | y1 | z1 (low bits first) inside it: rows of 4 along x
    are contiguous (a cell's x-neighbour corners are adjacent entries unless
    x is the block's last), and a cell whose corner 0 has even coordinates
    keeps its 8 corners in one 16-entry group"""
    T = 1 << 19
    idx = lambda x, y, z: oracle.lib.ivros_index(4, 16, 1.5, T, x, y, z)  # noqa: E731
    slot = (1 ^ 2654435761 ^ 805459861) & (T // 64 - 1)
    assert idx(5, 6, 7) == slot * 64 + 57  # x=1 (bits 0-1), y0=0, z0=1 (bit 3), y1=1 (bit 4), z1=1 (bit 5)
    assert idx(4, 4, 4) == slot * 64
    block = sorted(idx(4 + a, 8 + b, 12 + c) for a in range(4) for b in range(4) for c in range(4))
    base = ((1 ^ 2 * 2654435761 ^ 3 * 805459861) & (T // 64 - 1)) * 64
    assert block == list(range(base, base + 64))
    for x in range(4, 7):  # x-pairs inside a block are adjacent entries, in order
        for y in range(8, 12):
            for z in range(12, 16):
                assert idx(x + 1, y, z) == idx(x, y, z) + 1
    cell = sorted(idx(6 + (c & 1), 10 + (c >> 1 & 1), 14 + (c >> 2)) for c in range(8))
    assert cell[0] % 16 == 2 and cell[-1] - cell[0] < 16 and cell == sorted(set(cell))
    # the dense levels keep x-pairs adjacent too: x + n1 (y + n1 z) with x < N
    assert oracle.lib.ivros_index(2, 16, 1.5, T, 36, 5, 7) == oracle.lib.ivros_index(2, 16, 1.5, T, 35, 5, 7) + 1


def test_mlp_and_encoding_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(rng)
    table = rng.uniform(-0.5, 0.5, (16, T_SMALL, 2)).astype(np.float32)
    xyz = rng.uniform(0.02, 0.98, (6, 3))
    t = rng.uniform(0, 1, 6).astype(np.float32)
    y, loss, g, gt = oracle.sg_forward_backward(p, table, T_SMALL, xyz, t, bf16=False)

    def L(pp, tb):
        return oracle.sg_forward_backward(pp, tb, T_SMALL, xyz, t, bf16=False)[1]
    # parameters in float32 storage: central differences with h = 2^-10
    h = 2.0 ** -10
    for idx in list(rng.choice(64 * 32, 4)) + [64 * 32 + 3, 2112 + 7 * 64 + 5, 6208 + 9, 6272 + 11, 6336]:
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (L(pp, table) - L(pm, table)) / (2 * h)
        assert abs(fd - g[idx]) <= 2e-3 * max(1.0, abs(fd)), (idx, fd, g[idx])
    # table entries touched by the first point at a dense and a hashed level;
    # a smaller step there keeps the perturbed encodings off the ReLU kinks
    h = 2.0 ** -13
    for lvl in (1, 9):
        nz = np.argwhere(np.abs(gt[lvl]) > 0)
        e, f = nz[0]
        tp, tm = table.copy(), table.copy()
        tp[lvl, e, f] += h
        tm[lvl, e, f] -= h
        fd = (L(p, tp) - L(p, tm)) / (2 * h)
        assert abs(fd - gt[lvl, e, f]) <= 2e-3 * max(1e-2, abs(fd)), (lvl, fd, gt[lvl, e, f])


def test_forward_only_equals_forward_of_training_path():
    """sg_eval (the CPU baseline's forward-only evaluation) gives exactly the
    outputs of the forward pass inside sg_forward_backward"""
    rng = np.random.default_rng(7)
    table = rng.normal(0.0, 0.1, (16, T_SMALL, 2)).astype(np.float32)
    p = _params(rng)
    xyz = rng.uniform(0.0, 1.0, (40, 3))
    y, _, _, _ = oracle.sg_forward_backward(p, table, T_SMALL, xyz, np.zeros(40, np.float32))
    assert np.array_equal(oracle.sg_eval(p, table, T_SMALL, xyz), y)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(2)
    p = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    z = np.zeros(100, np.float32)
    p1, m1, v1 = oracle.sg_adam(p, g, z, z, 1e-2, 0.9, 0.99, 1e-15, 1)
    # bias-corrected first step moves every parameter by lr * sign(g)
    assert np.allclose(p1, p - 1e-2 * np.sign(g), atol=1e-6)
    assert np.allclose(m1, 0.1 * g, rtol=1e-6) and np.allclose(v1, 0.01 * g * g, rtol=1e-5)


def _corner_coords(T, l, n1, hashed):
    """per table entry of level l, the vertex coordinates it holds: dense
    entries decode x + n1 (y + n1 z); a hashed entry's place in its 64-entry
    slot decodes x1 x0 | y0 | z0 | y1 | z1 (low bits first) to the vertex's
    coordinates inside its 4x4x4 block"""
    e = np.arange(T, dtype=np.int64)
    if not hashed:
        return e % n1, (e // n1) % n1, e // (n1 * n1)
    loc = e & 63
    bit = lambda k: (loc >> k) & 1  # noqa: E731
    return loc & 3, bit(2) | bit(4) << 1, bit(3) | bit(5) << 1


def test_encoding_reproduces_affine_fields():
    """a level whose entries hold an affine function of their vertex's
    coordinates encodes every point as that function of x * N (dense levels:
    global coordinates; hashed levels: coordinates inside the 4x4x4 block,
    for cells inside one block, whatever slots the blocks share); a wrong
    corner weight, axis or place bit breaks it"""
    T = 1 << 17  # levels 0-2 dense (17^3, 25^3, 37^3 entries), 3-15 hashed
    rng = np.random.default_rng(5)
    table = np.zeros((16, T, 2), np.float32)
    coef = rng.integers(-64, 65, (16, 2, 4)) / 64.0  # a.(x, y, z) + c per feature, exact in float32 entries
    hashed = []
    for l in range(16):
        N = oracle.lib.ivros_level_res(l, 16, 1.5)
        h = (N + 1) ** 3 > T
        hashed.append(h)
        x, y, z = _corner_coords(T, l, N + 1, h)
        for k in range(2):
            a = coef[l, k]
            table[l, :, k] = (a[0] * x + a[1] * y + a[2] * z + a[3]).astype(np.float32)
    assert hashed[:3] == [False] * 3 and all(hashed[3:])
    checked = [0] * 16
    for x in rng.uniform(0.0, 1.0, (300, 3)):
        enc = oracle.sg_encode(table, T, x)
        for l in range(16):
            N = oracle.lib.ivros_level_res(l, 16, 1.5)
            u = x * N
            v0 = np.minimum(np.floor(u).astype(int), N - 1)
            if hashed[l]:
                if ((v0 & 3) == 3).any():
                    continue  # the cell spans two blocks
                u = u - 4 * (v0 // 4)
            for k in range(2):
                a = coef[l, k]
                assert abs(enc[2 * l + k] - (a[:3] @ u + a[3])) < 1e-9, (l, k, x)
            checked[l] += 1
    assert min(checked) > 50


def _mlp64(p, e):
    """W3 relu(W2 relu(W1 e + b1) + b2) + b3 in float64, parameters in the
    W1 b1 W2 b2 W3 b3 order"""
    p = p.astype(np.float64)
    W1, b1 = p[:2048].reshape(64, 32), p[2048:2112]
    W2, b2 = p[2112:6208].reshape(64, 64), p[6208:6272]
    W3, b3 = p[6272:6336], p[6336]
    h1 = np.maximum(W1 @ e + b1, 0.0)
    h2 = np.maximum(W2 @ h1 + b2, 0.0)
    return W3 @ h2 + b3


def test_mlp_forward_is_the_textbook_mlp():
    """the oracle's double-precision forward (encoding then MLP) equals
    numpy's W3 relu(W2 relu(W1 x + b1) + b2) + b3 on its encodings; with the
    bf16 rounding points on it stays within bf16 precision of it"""
    rng = np.random.default_rng(8)
    table = rng.normal(0.0, 0.2, (16, T_SMALL, 2)).astype(np.float32)
    p = _params(rng)
    xyz = rng.uniform(0.0, 1.0, (64, 3))
    ref = np.array([_mlp64(p, oracle.sg_encode(table, T_SMALL, x)) for x in xyz])
    y64 = oracle.sg_eval(p, table, T_SMALL, xyz, bf16=False)
    assert np.abs(y64 - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    y16 = oracle.sg_eval(p, table, T_SMALL, xyz, bf16=True)
    assert np.abs(y16 - ref).max() < 3e-2 * max(1.0, np.abs(ref).max())
    assert np.abs(y16 - ref).max() > 0.0  # the rounding points are really on


def test_psnr_ssim_closed_forms():
    """PSNR of an image against itself offset by d is 10 log10(255^2 / d^2);
    SSIM of an image with itself is 1, and of two constant images a and b it
    is (2ab + C1) / (a^2 + b^2 + C1)"""
    rng = np.random.default_rng(9)
    img = rng.integers(10, 200, (40, 48, 4)).astype(np.uint8)
    for d in (1, 3, 17):
        assert abs(ivr.psnr(img + np.uint8(d), img) - 10.0 * np.log10(255.0 ** 2 / d ** 2)) < 1e-9
    assert ivr.psnr(img, img) == 99.0
    assert abs(ivr.ssim(img, img) - 1.0) < 1e-12
    a, b = 80.0, 120.0
    ia = np.full((30, 30, 4), a, np.uint8)
    ib = np.full((30, 30, 4), b, np.uint8)
    c1 = (0.01 * 255) ** 2
    # luminance weights sum to 1, so a constant RGB image has luminance a
    assert abs(ivr.ssim(ia, ib) - (2 * a * b + c1) / (a * a + b * b + c1)) < 1e-9


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_render_network_value_matches_oracle(ctx, small_vol):
    """the network value the renderer composites (ivr_surrogate_eval_points:
    the device function of k_render_sg, half-precision tables, bf16 MFMA)
    equals the oracle's forward pass on the fp16-rounded tables, within the
    bf16 tolerance of the training-path check, at points of several bricks"""
    vol, _ = small_vol
    sg = ctx.surrogate(vol, log2_table=12)
    rng = np.random.default_rng(31)
    p = _params(rng)
    tables = [rng.uniform(-0.5, 0.5, (16, T_SMALL, 2)).astype(np.float32) for _ in range(3)]
    sg.set(mlp=p, table=tables[0], brick=0)
    sg.set(table=tables[1], brick=3)
    sg.set(table=tables[2], brick=7)
    n = 300
    xyz = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    xyz[:8] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1, 0, 0.25], [0, 1, 0.75], [0.999, 0.001, 0.5],
               [0.25, 0.75, 1.0], [1 / 3, 2 / 3, 0.1]]
    bricks = np.array([0, 3, 7], np.int32)[rng.integers(0, 3, n)]
    y = sg.eval_points(xyz, bricks)
    yo = np.zeros(n)
    for b, t in zip((0, 3, 7), tables):
        sel = bricks == b
        yo[sel] = oracle.sg_eval(p, t.astype(np.float16).astype(np.float32), T_SMALL, xyz[sel].astype(np.float64))
    err = np.abs(y - yo).max()
    print("max |y_render - y_oracle| = %.3e (|y| up to %.3f)" % (err, np.abs(yo).max()))
    assert err < 2e-3 * max(1.0, np.abs(yo).max())
    # the fp32 tables would differ from the half copy: the check sees which one the renderer reads
    yf = np.zeros(n)
    for b, t in zip((0, 3, 7), tables):
        sel = bricks == b
        yf[sel] = oracle.sg_eval(p, t, T_SMALL, xyz[sel].astype(np.float64))
    assert np.abs(y - yo).mean() < np.abs(y - yf).mean()
    sg.close()

@pytest.fixture(scope="module")
def small_vol(ctx):
    gp = configs.gen_params(2)
    host = ivr.generate_host(gp, 64, 64, 64)
    vol = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=32)
    vol.upload(host)
    yield vol, host
    vol.close()


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["scattered", "clustered"])
def test_gpu_gradients_match_oracle(ctx, small_vol, layout):
    """GPU gradients of explicit samples against the oracle.  Scattered:
    uniform points of two bricks, every level's table gradient added by
    direct atomics.  Clustered: each wavefront's 64 points in one brick and a
    cube of edge 0.05, so the coarse levels' gradients are summed in LDS
    before they reach the table (and the finer ones are not)."""
    vol, _ = small_vol
    sg = ctx.surrogate(vol, log2_table=12)
    rng = np.random.default_rng(3)
    p = _params(rng)
    tables = [rng.uniform(-0.5, 0.5, (16, T_SMALL, 2)).astype(np.float32) for _ in range(2)]
    sg.set(mlp=p, table=tables[0], brick=0)
    sg.set(table=tables[1], brick=1)
    n = 200  # not a multiple of 64: exercises the padded tile
    if layout == "scattered":
        xyz = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
        bricks = rng.integers(0, 2, n).astype(np.int32)
    else:
        groups = (n + 63) // 64
        corner = rng.uniform(0.0, 0.95, (groups, 3))
        xyz = (np.repeat(corner, 64, 0)[:n] + rng.uniform(0.0, 0.05, (n, 3))).astype(np.float32)
        bricks = np.repeat(np.arange(groups) % 2, 64)[:n].astype(np.int32)
    t = rng.uniform(0, 1, n).astype(np.float32)
    y, loss, g, gt0 = sg.grad(xyz, bricks, t, table_brick=0)
    # the training forward pass reads the half-precision copy of the tables
    tables = [tb.astype(np.float16).astype(np.float32) for tb in tables]
    # oracle: both bricks' points, MLP gradient summed; table gradient of brick 0
    yo = np.zeros(n)
    go = np.zeros(oracle.SG_NPARAM)
    lo = 0.0
    gto = None
    for b in (0, 1):
        sel = bricks == b
        yb, lb, gb, gtb = oracle.sg_forward_backward(p, tables[b], T_SMALL, xyz[sel].astype(np.float64), t[sel])
        yo[sel] = yb
        nb = sel.sum()
        go += gb * nb / n  # oracle normalises by its own batch size
        lo += lb * nb / n
        if b == 0:
            gto = gtb * nb / n
    print("max |y - y_oracle| = %.3e, loss %.6f vs %.6f" % (np.abs(y - yo).max(), loss, lo))
    assert np.abs(y - yo).max() < 2e-3 * max(1.0, np.abs(yo).max())
    assert abs(loss - lo) < 1e-3 * max(lo, 1e-3)
    rel = np.abs(g - go).max() / np.abs(go).max()
    l2 = lambda a: float(np.sqrt(np.sum(np.square(a, dtype=np.float64))))
    relt = l2(gt0 - gto) / l2(gto)
    # the bf16 rounding noise of the same batch: bf16-emulating vs pure double oracle
    sel = bricks == 0
    _, _, _, gt64 = oracle.sg_forward_backward(p, tables[0], T_SMALL, xyz[sel].astype(np.float64), t[sel],
                                               bf16=False)
    noise = l2(gto - gt64 * sel.sum() / n) / l2(gto)
    per_level = [l2(gt0[k] - gto[k]) / max(l2(gto[k]), 1e-30) for k in range(16)]
    print("MLP gradient max rel err %.3e; table gradient rel L2 err %.3e (bf16 noise %.3e, max entry err %.3e)"
          % (rel, relt, noise, np.abs(gt0 - gto).max() / np.abs(gto).max()))
    print("table gradient rel L2 err per level:", " ".join("%.1e" % e for e in per_level))
    assert rel < 2e-2
    assert relt < max(2e-2, 3.0 * noise)
    # a wrong index or level would decorrelate whole levels
    assert max(per_level) < 0.25
    # Adam on the GPU against the oracle's step from the same gradients
    sg.adam()
    p1, _ = sg.get(table=False)
    z = np.zeros_like(p)
    po, _, _ = oracle.sg_adam(p, g, z, z, sg.sp.lr_mlp, sg.sp.beta1, sg.sp.beta2, sg.sp.eps, 1)
    assert np.abs(p1 - po).max() < 1e-6 + 1e-5 * np.abs(po).max()
    sg.close()


def _write_morton_checkpoint(src, dst, sg):
    """rewrite checkpoint src (format 5) as format 3: the same sections with
    every hashed level's slots in Morton order and macrocell ranges over
    16^3-voxel cells, the layout of that format"""
    import ctypes
    raw = bytearray(open(src, "rb").read())
    assert raw[:8] == b"IVRSG005"
    hdr = np.frombuffer(raw, np.int32, 8, 8 + ctypes.sizeof(ivr.SurrogateParams))
    off = 8 + ctypes.sizeof(ivr.SurrogateParams) + 8 * 4
    nparam = 6337
    ntab = sg.nbricks * 16 * sg.T * 2
    to = np.zeros(64, np.int64)  # Morton place -> place in the current order
    for m in range(64):
        x = (m & 1) | (m >> 2 & 2)
        y = (m >> 1 & 1) | (m >> 3 & 2)
        z = (m >> 2 & 1) | (m >> 4 & 2)
        to[m] = x | (y & 1) << 2 | (z & 1) << 3 | (y & 2) << 3 | (z & 2) << 4
    hashed = [(r + 1) ** 3 > sg.T for r in sg.res]
    for k in range(3):  # table values, first and second moments
        o = off + 4 * (3 * nparam + k * ntab)
        t = np.frombuffer(raw, np.float32, ntab, o).reshape(sg.nbricks, 16, sg.T // 64, 64, 2).copy()
        for lvl in range(16):
            if hashed[lvl]:
                t[:, lvl] = t[:, lvl][:, :, to, :]
        raw[o:o + 4 * ntab] = t.tobytes()
    ranges = off + 4 * (3 * nparam + 3 * ntab)
    n4 = int(np.prod([((int(n) - 2) >> 2) + 1 for n in hdr[2:5]]))
    n16 = int(np.prod([((int(n) - 2) >> 4) + 1 for n in hdr[2:5]]))
    assert len(raw) == ranges + 4 * n4  # a half (min, max) pair per 4^3 cell
    # ranges of 16^3 cells (the full range [0, 1] each): the loader reads and drops them
    raw = raw[:ranges] + np.tile(np.float32([0.0, 1.0]), n16).tobytes()
    raw[:8] = b"IVRSG003"
    open(dst, "wb").write(bytes(raw))


@pytest.mark.gpu
def test_training_gain_fallback_and_checkpoint(ctx, small_vol):
    vol, host = small_vol
    lut = configs.tf_lut(2)
    tf = ctx.tf(lut)
    st = configs.state(2, 160, 144)
    d = ivr.config_desc(st)
    fr = ctx.frame(160, 144)
    ctx.render(vol, tf, d, fr)
    plain = fr.download()
    sg = ctx.surrogate(vol, log2_table=15)
    sg.render(vol, tf, d, fr)
    untrained = fr.download()
    losses = [sg.train(vol, 40, 1 << 18) for _ in range(5)]
    sg.render(vol, tf, d, fr)
    trained = fr.download()
    p0, p1 = ivr.psnr(untrained, plain), ivr.psnr(trained, plain)
    print("losses", losses, "PSNR untrained %.2f dB trained %.2f dB SSIM %.4f" % (p0, p1, ivr.ssim(trained, plain)))
    assert losses[-1] < losses[0]
    assert p1 > p0 + 10.0
    # macrocell skipping: the network only sees samples that can contribute
    ds = ivr.config_desc(st, skip=True)
    sg.render(vol, tf, ds, fr)
    skipped = fr.download()
    p2 = ivr.psnr(skipped, plain)
    print("PSNR with macrocell skipping %.2f dB" % p2)
    assert p2 > p1 - 1.0
    # all bricks forced to fall back: identical to the plain render, with and without skipping
    sg.set_fallback(np.ones(sg.nbricks, np.uint8))
    sg.render(vol, tf, d, fr)
    assert np.array_equal(fr.download(), plain)
    sg.render(vol, tf, ds, fr)
    assert np.array_equal(fr.download(), plain)
    # per-brick evaluation at an unreachable threshold flags every brick
    psnr, nf = sg.evaluate(vol, tf, d, threshold_db=200.0)
    assert nf == sg.nbricks and (psnr > 0).all()
    psnr, nf = sg.evaluate(vol, tf, d, threshold_db=0.0)
    assert nf == 0
    # checkpoint round trip reproduces the render
    sg.render(vol, tf, d, fr)
    before = fr.download()
    path = os.path.join(tempfile.mkdtemp(), "sg.ckpt")
    sg.save(path)
    sg2 = ctx.surrogate(vol, log2_table=15, seed=99)
    sg2.load(path)
    sg2.render(vol, tf, d, fr)
    assert np.array_equal(fr.download(), before)
    sg2.close()
    # a file of the previous format (hashed slots in Morton order, x0 y0 z0 x1 y1 z1)
    # loads into the current order: the same render
    old = os.path.join(os.path.dirname(path), "sg3.ckpt")
    _write_morton_checkpoint(path, old, sg)
    sg3 = ctx.surrogate(vol, log2_table=15, seed=7)
    sg3.load(old)
    sg3.render(vol, tf, d, fr)
    assert np.array_equal(fr.download(), before)
    sg3.close()
    sg.close()
    fr.close()
    tf.close()


@pytest.mark.gpu
def test_cell_ranges_contain_the_field(ctx):
    """The surrogate's macrocell ranges are half pairs rounded outward.  A
    field valued exactly f32(42/255) classifies into LUT entry 42 (visible),
    while the nearest half, 0.16467..., falls into entry 41 (transparent):
    ranges rounded to nearest would mark every cell of the block transparent
    and skip it.  With every brick falling back, the surrogate's skipping
    frame must equal the plain frame."""
    v = np.float32(42) / np.float32(255)
    assert int(v * np.float32(255)) == 42
    assert int(np.float32(np.float16(v)) * np.float32(255)) == 41
    host = np.zeros((64, 64, 64), np.float32)
    host[16:48, 12:44, 20:52] = v
    vol = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=32)
    vol.upload(host)
    lut = np.zeros((256, 4), np.float32)
    lut[42:] = (1.0, 0.5, 0.2, 0.3)
    tf = ctx.tf(lut)
    st = configs.state(2, 160, 144)
    fr = ctx.frame(160, 144)
    ctx.render(vol, tf, ivr.config_desc(st), fr)
    plain = fr.download()
    assert plain[..., 3].max() > 0  # the block is visible
    sg = ctx.surrogate(vol, log2_table=12)
    sg.set_fallback(np.ones(sg.nbricks, np.uint8))
    sg.render(vol, tf, ivr.config_desc(st, skip=True), fr)
    assert np.array_equal(fr.download(), plain)
    sg.close()
    fr.close()
    tf.close()
    vol.close()


@pytest.mark.gpu
def test_gradient_over_several_chunks(ctx, small_vol):
    """Training runs in chunks of 2^22 samples (network pass, then the table
    scatter on a second stream, encoding gradients double-buffered).  Over
    2^22 + 2^20 + 37 samples, two chunks and the second ragged, the network
    outputs equal those of the two parts computed alone, bit for bit, and
    the mean gradients equal the parts' gradients weighted by their sample
    counts, up to the rounding of float atomics summed in another order."""
    vol, _ = small_vol
    sg = ctx.surrogate(vol, log2_table=15, seed=5)
    rng = np.random.default_rng(12)
    n1, n2 = 1 << 22, (1 << 20) + 37
    n = n1 + n2
    xyz = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    bricks = rng.integers(0, sg.nbricks, n).astype(np.int32)
    t = rng.uniform(0.0, 1.0, n).astype(np.float32)
    y, loss, g, gt = sg.grad(xyz, bricks, t, table_brick=3)
    y1, l1, g1, gt1 = sg.grad(xyz[:n1], bricks[:n1], t[:n1], table_brick=3)
    y2, l2, g2, gt2 = sg.grad(xyz[n1:], bricks[n1:], t[n1:], table_brick=3)
    assert np.array_equal(y, np.concatenate([y1, y2]))
    assert abs(loss - (n1 * l1 + n2 * l2) / n) <= 1e-5 * loss
    gw = (n1 * g1.astype(np.float64) + n2 * g2) / n
    assert np.abs(g - gw).max() <= 1e-4 * np.abs(gw).max()
    # a coarse vertex sums about a thousand per-sample terms of ~1e-7 that
    # largely cancel, so fp32 rounding of the partial sums reaches ~1e-3 of the
    # largest entry; a chunk lost or added, or read from the wrong buffer,
    # moves entries by tens of percent
    tw = (n1 * gt1.astype(np.float64) + n2 * gt2) / n
    assert np.abs(gt - tw).max() <= 1e-2 * np.abs(tw).max()
    assert np.linalg.norm(gt - tw) <= 1e-2 * np.linalg.norm(tw)
    assert np.count_nonzero(gt) > 0
    sg.close()


@pytest.mark.gpu
def test_helper_lanes_change_nothing(ctx, small_vol):
    """Lanes whose ray has ended evaluate live rays' next samples: the same
    trained network, loaded into a context made with IVR_SG_HELPERS=0, gives
    the same frames, sample counts and per-brick evaluation, with and
    without skipping and with some bricks falling back."""
    vol, host = small_vol
    lut = configs.tf_lut(2)
    st = configs.state(2, 160, 144)
    tf = ctx.tf(lut)
    fr = ctx.frame(160, 144)
    sg = ctx.surrogate(vol, log2_table=15)
    for _ in range(3):
        sg.train(vol, 20, 1 << 18)
    path = os.path.join(tempfile.mkdtemp(), "sg.ckpt")
    sg.save(path)
    os.environ["IVR_SG_HELPERS"] = "0"
    try:
        ctx2 = ivr.Context(0)
    finally:
        del os.environ["IVR_SG_HELPERS"]
    vol2 = ctx2.volume(64, 64, 64, ivr.IVR_F32, brick=32)
    vol2.upload(host)
    tf2 = ctx2.tf(lut)
    fr2 = ctx2.frame(160, 144)
    sg2 = ctx2.surrogate(vol2, log2_table=15)
    sg2.load(path)
    fallback = (np.arange(sg.nbricks) % 3 == 1).astype(np.uint8)
    for flags in (np.zeros(sg.nbricks, np.uint8), fallback):
        sg.set_fallback(flags)
        sg2.set_fallback(flags)
        for skip in (False, True):
            d = ivr.config_desc(st, skip=skip)
            sg.render(vol, tf, d, fr)
            sg2.render(vol2, tf2, d, fr2)
            a, b = fr.download(), fr2.download()
            assert a[..., 3].max() > 0
            assert np.array_equal(a, b), (flags.tolist(), skip)
            sa, sb = fr.stats(), fr2.stats()
            assert (sa["samples"], sa["visited"]) == (sb["samples"], sb["visited"])
            assert sa["evaluations"] <= sb["evaluations"]  # helpers fill lane slots
    d = ivr.config_desc(st, skip=True)
    pa, na = sg.evaluate(vol, tf, d, threshold_db=40.0)
    pb, nb = sg2.evaluate(vol2, tf2, d, threshold_db=40.0)
    assert na == nb and np.array_equal(pa, pb)
    for o in (sg2, fr2, tf2, vol2, sg, fr, tf):
        o.close()
    ctx2.close()


@pytest.mark.gpu
def test_grouped_bricks_and_freeze(ctx, small_vol):
    """A surrogate whose bricks span 2x2x2 volume bricks (log2_group=1) over the
    64^3 volume in 32^3 bricks is the surrogate of the same volume in one 64^3
    brick: the same training samples (equal first-step loss from equal
    initial weights) and, with the same weights, bit-identical frames.  Forced
    fallback through the grouped geometry equals the plain render.  Freezing
    releases 16 bytes per table parameter and the training buffers, and changes
    no frame."""
    vol32, host = small_vol
    vol64 = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=64)
    vol64.upload(host)
    sg_g = ctx.surrogate(vol32, log2_table=12, log2_group=1, seed=77)
    sg_1 = ctx.surrogate(vol64, log2_table=12, seed=77)
    assert sg_g.nbricks == 1 and sg_1.nbricks == 1
    with pytest.raises(ivr.IvrError):
        sg_g.train(vol64, 1, 1 << 12)  # a volume of other bricks than the surrogate was made for
    lg, l1 = sg_g.train(vol32, 1, 1 << 16), sg_1.train(vol64, 1, 1 << 16)
    print("first-step loss grouped %.6e, one brick %.6e" % (lg, l1))
    assert abs(lg - l1) <= 1e-4 * l1
    sg_g.train(vol32, 20, 1 << 16)
    mlp, tb = sg_g.get(0)
    sg_1.set(mlp=mlp, table=tb, brick=0)
    tf = ctx.tf(configs.tf_lut(2))
    st = configs.state(2, 160, 144)
    fr = ctx.frame(160, 144)
    frames = {}
    for skip in (False, True):
        d = ivr.config_desc(st, skip=skip)
        sg_g.render(vol32, tf, d, fr)
        a = fr.download()
        sg_1.render(vol64, tf, d, fr)
        assert np.array_equal(a, fr.download())
        frames[skip] = a
    d = ivr.config_desc(st)
    ctx.render(vol32, tf, d, fr)
    plain = fr.download()
    print("grouped surrogate PSNR %.2f dB" % ivr.psnr(frames[False], plain))
    sg_g.set_fallback(np.ones(1, np.uint8))
    sg_g.render(vol32, tf, d, fr)
    assert np.array_equal(fr.download(), plain)
    sg_g.set_fallback(np.zeros(1, np.uint8))
    # freeze: the training state goes, the frames stay
    xyz = np.random.default_rng(5).uniform(0, 1, (100, 3)).astype(np.float32)
    y0 = sg_g.eval_points(xyz, np.zeros(100, np.int32))
    b0 = ctx.device_bytes()
    sg_g.freeze()
    freed = b0 - ctx.device_bytes()
    # tables, gradients and moments (16 bytes per parameter), the MLP's gradient and
    # moments, and the encoding-gradient buffers of two training chunks (2^16 samples x 32 fp32)
    assert freed == sg_g.nbricks * 16 * sg_g.T * 2 * 16 + 3 * ivr.SG_NPARAM * 4 + 2 * (1 << 16) * 32 * 4
    sg_g.render(vol32, tf, d, fr)
    assert np.array_equal(fr.download(), frames[False])
    sg_g.render(None, tf, d, fr)  # no plain volume resident
    assert np.array_equal(fr.download(), frames[False])
    assert np.array_equal(sg_g.eval_points(xyz, np.zeros(100, np.int32)), y0)
    for call in (lambda: sg_g.train(vol32, 1, 1 << 12), lambda: sg_g.get(0),
                 lambda: sg_g.save(os.path.join(tempfile.mkdtemp(), "x"))):
        with pytest.raises(ivr.IvrError):
            call()
    sg_g.freeze()  # a second freeze is a no-op
    # a checkpoint of the one-brick surrogate does not load into the grouped one
    path = os.path.join(tempfile.mkdtemp(), "sg1.ckpt")
    sg_1.save(path)
    sg_h = ctx.surrogate(vol32, log2_table=12, log2_group=1)
    with pytest.raises(ivr.IvrError):
        sg_h.load(path)
    for o in (sg_h, sg_g, sg_1, fr, tf, vol64):
        o.close()
