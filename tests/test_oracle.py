"""Pins the CPU oracle (oracle/ivr_oracle.c).

Compat mode, sampler, slab test, sample marching, transfer functions,
colormaps and tile compositing are compared bit for bit with golden vectors
written by the reference renderer compiled from its own sources
(tests/golden/make_golden.py).  Config mode, which has no reference
counterpart, is pinned by closed forms: true trilinear reproduces affine
fields, a constant slab gives A = 1-(1-a)^n and C = c A, early termination
drops at most 1-A_stop of the remaining radiance, back-to-front and
front-to-back compositing agree, and the bricked (tile) decomposition
composited with "over" equals the unbricked render.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import ivr_amd as ivr
import oracle
from ivr_amd import configs

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(G, "golden_meta.json")))


def f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@pytest.fixture(scope="module")
def vol255():
    gp = configs.gen_params(1)
    gp.value_scale = 255.0
    v = ivr.generate_host(gp, 64, 64, 64)
    assert hashlib.sha256(v.tobytes()).hexdigest() == META["sha256"]
    return v


def test_box_matches_reference():
    bmin, bmax = oracle.box(64, 64, 64)
    assert np.array_equal(np.concatenate([bmin, bmax]), np.array(META["box"], np.float32))
    bmin, bmax = oracle.box(501, 201, 276)  # the reference's shell dataset dims
    assert np.allclose(bmax, [1.0, 201 / 501, 276 / 501], rtol=1e-6) and np.array_equal(bmin, -bmax)


def test_compat_images_bit_exact(vol255):
    z = np.load(os.path.join(G, "compat_images.npz"))
    luts = {"tf5": z["lut_tf5"], "medical": z["lut_medical"], "jet": z["lut_jet"]}
    for k, case in enumerate(META["compat_cases"]):
        st = ivr.state_from_angles(case["az"], case["el"], META["width"], META["height"])
        st.alpha = case["alpha"]
        p = oracle.params(mode=oracle.COMPAT, xfer=case["xfer"], lutscale=1.0,
                          maxlutindx=case["maxlutindx"])
        img, n = oracle.render(vol255, oracle.view_of(st), luts[case["lut"]], p)
        ref = z["images"][k]
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), case
        assert n > 0


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


def test_compat_cut_planes_pan_and_wide_rect_bit_exact(vol255):
    """compat frames with STATE cut planes (dconst = min(box, newconst),
    ray.c:44-51), a panned eye (buildray, mathutil.c:1129-1137) and a box
    projected past the viewport (the pixel rectangle's one-row overflow,
    main.c:959-962), against the compiled reference"""
    z = np.load(os.path.join(G, "compat_cuts.npz"))
    W, H = META["width"], META["height"]
    for k, row in enumerate(z["states"]):
        st = _state_of(row, W, H)
        p = oracle.params(mode=oracle.COMPAT, xfer=3, lutscale=1.0, maxlutindx=255)
        img, n = oracle.render(vol255, oracle.view_of(st), z["lut"], p)
        assert np.array_equal(img.view(np.uint32), z["images"][k].view(np.uint32)), k
        assert n > 0


def test_sampler_bit_exact(vol255):
    z = np.load(os.path.join(G, "sampler.npz"))
    bmin, bmax = oracle.box(64, 64, 64)
    got = np.array([oracle.lib.ivro_sample_ref(vol255.ctypes.data_as(C.c_void_p), 64, 64, 64,
                                               bmin.ctypes.data_as(C.c_void_p),
                                               bmax.ctypes.data_as(C.c_void_p), *p)
                    for p in z["points"]], np.float32)
    assert np.array_equal(got.view(np.uint32), z["values"].view(np.uint32))


def test_nearest_sampler_bit_exact(vol255):
    """getnearestsample (ray.c:182-238): the voxel's bits as an unsigned int,
    as float, at 200 points against the compiled reference"""
    z = np.load(os.path.join(G, "nearest.npz"))
    bmin, bmax = oracle.box(64, 64, 64)
    vals = np.array([oracle.lib.ivro_sample_nearest(vol255.ctypes.data_as(C.c_void_p), 64, 64, 64,
                                                     bmin.ctypes.data_as(C.c_void_p),
                                                     bmax.ctypes.data_as(C.c_void_p), *map(float, p))
                     for p in z["points"]], np.float32)
    assert np.array_equal(vals.view(np.uint32), z["values"].view(np.uint32))


def test_slab_bit_exact():
    z = np.load(os.path.join(G, "slab.npz"))
    bmin, bmax = oracle.box(64, 64, 64)
    nc = np.full(6, 10.0, np.float32)
    for r, hit, tin, tout in zip(z["rays"], z["hit"], z["tin"], z["tout"]):
        ti, to = C.c_float(), C.c_float()
        h = oracle.lib.ivro_rayintersect(f(bmin), f(bmax), f(nc), f(np.ascontiguousarray(r)),
                                         C.byref(ti), C.byref(to))
        assert h == hit
        assert np.float32(ti.value).tobytes() == np.float32(tin).tobytes()
        assert np.float32(to.value).tobytes() == np.float32(tout).tobytes()
    assert z["hit"].sum() > 50 and (z["hit"] == 0).sum() > 20


def test_samplesonray_bit_exact(vol255):
    z = np.load(os.path.join(G, "samplesonray.npz"))
    nc = np.full(6, 10.0, np.float32)
    for r, n, s in zip(z["rays"], z["n"], z["samples"]):
        buf = np.zeros(256, np.float32)
        got = oracle.lib.ivro_samplesonray_compat(vol255.ctypes.data_as(C.c_void_p), 64, 64, 64,
                                                  f(nc), f(np.ascontiguousarray(r)), 256, f(buf))
        assert got == n
        if n > 0:
            assert np.array_equal(buf[:n].view(np.uint32), s[:n].view(np.uint32))


def test_voxel_walk_bit_exact():
    """getfirstvoxel / getvoxelsonray (ray.c:333-483) against the compiled
    reference: 120 rays (world and cell-coordinate origins, zero direction
    components, face ties from grid-point origins) at the volume's dims and
    at non-cubic dims"""
    z = np.load(os.path.join(G, "dda.npz"))
    md = int(z["maxdepth"])
    for k, r in enumerate(z["rays"]):
        rc, v = oracle.firstvoxel(z["box"], r)
        assert rc == z["rc"][k]
        if rc == 0:
            assert v == list(z["first"][k]), k
        for tag in ("own", "odd"):
            rc, w = oracle.voxelsonray(z["box"], z["dims_" + tag], r, md)
            assert np.array_equal(w.view(np.uint32), z["voxels_" + tag][k].view(np.uint32)), (k, tag)


def test_voxel_walk_steps_are_face_neighbours():
    """every step of the walk moves to a face neighbour: the offset
    ((z*h + x)*w + y changes by exactly w (x), 1 (y) or w*h (z), in the
    direction's sign, and each axis advances at the rate 1/|d| predicts"""
    rng = np.random.default_rng(5)
    box = np.float32([-1, -1, -1, 1, 1, 1])
    w, h, d = 40, 56, 72
    for _ in range(50):
        o = rng.uniform(0.0, 30.0, 3).astype(np.float32)
        dr = rng.normal(size=3).astype(np.float32)
        rc, walk = oracle.voxelsonray(box, (w, h, d), np.concatenate([o, dr]), 200)
        assert rc == 0
        steps = np.diff(walk.astype(np.int64))
        allowed = {int(np.sign(dr[0])) * w, int(np.sign(dr[1])), int(np.sign(dr[2])) * w * h}
        assert set(steps.tolist()) <= allowed
        counts = np.array([np.sum(steps == int(np.sign(dr[0])) * w), np.sum(steps == int(np.sign(dr[1]))),
                           np.sum(steps == int(np.sign(dr[2])) * w * h)], np.float64)
        expect = 199 * np.abs(dr) / np.abs(dr).sum()
        assert np.all(np.abs(counts - expect) <= 0.02 * 199 * 3 + 3), (counts, expect)


def test_voxel_walk_zero_direction_misses():
    box = np.float32([-1, -1, -1, 1, 1, 1])
    rc, v = oracle.firstvoxel(box, np.float32([0.3, 0.2, 0.1, 0, 0, 0]))
    assert rc == -1
    rc, w = oracle.voxelsonray(box, (8, 8, 8), np.float32([0.3, 0.2, 0.1, 0, 0, 0]), 4)
    assert rc == -1 and np.all(w == -7.0)
    # along +x only: offsets advance by w per step
    rc, w = oracle.voxelsonray(box, (8, 9, 10), np.float32([2.25, 3.5, 4.5, 1.0, 0, 0]), 5)
    assert rc == 0 and np.all(np.diff(w) == 8)


def test_gradient_is_numpy_gradient():
    """the oracle's gradient field == numpy.gradient (central differences
    (f[i+1]-f[i-1])/2 inside, one-sided first differences on the faces, here
    halved as gradient3d divides every difference by 2), magnitude the float32
    sqrt of (x*x + y*y) + z*z: bit for bit, on a ragged random volume"""
    v = np.random.default_rng(1).normal(size=(7, 9, 11)).astype(np.float32)
    g = oracle.gradient(v)
    e = np.stack([np.gradient(v, axis=2), np.gradient(v, axis=1), np.gradient(v, axis=0)], -1)
    for comp, axis in ((0, 2), (1, 1), (2, 0)):
        for idx in (0, -1):
            sl = [slice(None)] * 3
            sl[axis] = idx
            e[tuple(sl) + (comp,)] *= np.float32(0.5)
    m = np.sqrt((e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1]) + e[..., 2] * e[..., 2])
    assert np.array_equal(g[..., :3].view(np.uint32), e.view(np.uint32))
    assert np.array_equal(g[..., 3].view(np.uint32), m.view(np.uint32))


def test_gradient_of_affine_field():
    """f = a*x + b*y + c*z + d on the voxel grid: (a, b, c) inside, half of
    the axis's slope on that axis's faces, the other components unchanged"""
    z, y, x = np.meshgrid(np.arange(6), np.arange(5), np.arange(8), indexing="ij")
    v = (0.5 * x - 0.25 * y + 2.0 * z + 1.0).astype(np.float32)
    g = oracle.gradient(v)
    assert np.all(g[1:-1, 1:-1, 1:-1, :3] == np.float32([0.5, -0.25, 2.0]))
    assert np.all(g[:, :, 0, 0] == 0.25) and np.all(g[:, :, -1, 0] == 0.25)
    assert np.all(g[:, 0, :, 1] == -0.125) and np.all(g[0, :, :, 2] == 1.0)
    assert np.all(g[1:-1, 1:-1, 0, 1:3] == np.float32([-0.25, 2.0]))
    assert np.allclose(g[2, 2, 2, 3], np.sqrt(0.25 + 0.0625 + 4.0))


def test_transfer_functions_bit_exact():
    z = np.load(os.path.join(G, "transferfuncs.npz"))
    lut = np.ascontiguousarray(z["lut"])
    for x in range(6):
        for k in range(12):
            o = np.zeros(4, np.float32)
            s = np.ascontiguousarray(z["samples"][k])
            oracle.lib.ivro_transferfunc(x, f(lut), int(z["lutoffset"]), int(z["maxlutindx"]), f(s),
                                         64 - k, C.c_float(float(z["alpha"])), f(o))
            assert np.array_equal(o.view(np.uint32), z["out"][x, k].view(np.uint32)), (x, k)


def test_colormaps_bit_exact():
    z = np.load(os.path.join(G, "colormaps.npz"))
    for key in z.files:
        cid, hsize = (int(t) for t in key[3:].split("_"))
        hmin, hmax = (0.0, 255.0) if hsize == 256 else (1468.0, 4985.0)
        out = np.zeros((hsize, 4), np.float32)
        oracle.lib.ivro_colormap(cid, hsize, C.c_float(hmin), C.c_float(hmax), f(out))
        assert np.array_equal(out.view(np.uint32), z[key].view(np.uint32)), key


def test_composite_tiles_bit_exact():
    z = np.load(os.path.join(G, "composite.npz"))
    got = oracle.composite_tiles(z["tiles"])
    assert np.array_equal(got.view(np.uint32), z["out"].view(np.uint32))


# ---------------------------------------------------------------- config mode closed forms

def _tri(vol, x, y, z):
    bmin, bmax = oracle.box(*vol.shape[::-1])
    return oracle.lib.ivro_sample_tri(vol.ctypes.data_as(C.c_void_p), vol.shape[2], vol.shape[1],
                                      vol.shape[0], bmin.ctypes.data_as(C.c_void_p),
                                      bmax.ctypes.data_as(C.c_void_p), x, y, z)


def test_trilinear_reproduces_affine_fields():
    n = 32
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    a, b, c, d = 0.25, 0.013, -0.007, 0.021
    vol = (a + b * i + c * j + d * k).astype(np.float32)  # i along x, j along y, k along z
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = rng.uniform(-0.95, 0.95, 3)
        g = (p + 1.0) * n / 2.0  # voxel coordinates: voxel v sits at min + v * ext / n
        g = np.minimum(g, n - 1)
        expect = a + b * g[0] + c * g[1] + d * g[2]
        assert abs(_tri(vol, *p) - expect) < 2e-6
    # a mistake that swaps x and y weights fails this field
    assert b != c


def test_constant_slab_closed_form():
    n = 16
    vol = np.full((n, n, n), 0.5, np.float32)
    lut = np.zeros((256, 4), np.float32)
    lut[127] = [0.8, 0.4, 0.2, 0.03]
    st = ivr.state_from_angles(0.0, 0.0, 8, 8)  # axis-aligned view along -z
    v = oracle.view_of(st)
    img, _ = oracle.render(vol, v, lut, oracle.params(ert=2.0))
    # the centre pixels' rays cross the box along z: n_s = (int)(1 + L/dt + 0.5)
    px = img[4, 4]
    A = px[3]
    ns = round(np.log(1.0 - A) / np.log(1.0 - 0.03))
    assert abs(A - (1.0 - (1.0 - 0.03) ** ns)) < 1e-5
    assert np.allclose(px[:3], lut[127, :3] * A, atol=1e-5)
    # a ray of the unit-cube box at step 0.5 voxel crosses ~2n samples
    assert 2 * n - 2 <= ns <= 2 * n + 3


def test_ert_residual_bound():
    gp = configs.gen_params(1)
    vol = ivr.generate_host(gp, 48, 48, 48)
    lut = configs.tf_lut(1)
    lut[:, 3] = np.minimum(lut[:, 3] * 20.0, 1.0)  # opaque enough to terminate
    st = ivr.state_from_angles(0.7, 0.3, 48, 48)
    on, n_on = oracle.render(vol, oracle.view_of(st), lut, oracle.params(ert=0.99))
    off, n_off = oracle.render(vol, oracle.view_of(st), lut, oracle.params(ert=2.0))
    assert n_on < n_off
    resid = off - on
    assert (resid >= -1e-6).all()
    assert (resid <= (1.0 - np.minimum(on[..., 3:], 1.0)) + 1e-6).all()


def test_btf_equals_ftb():
    rng = np.random.default_rng(3)
    lut = np.ascontiguousarray(rng.uniform(0, 1, (256, 4)).astype(np.float32))
    lut[:, 3] *= 0.3
    s = np.ascontiguousarray(rng.uniform(0, 255, 200).astype(np.float32))
    o = np.zeros(4, np.float32)
    oracle.lib.ivro_transferfunc(3, f(lut), 0, 0, f(s), 200, C.c_float(1.0), f(o))
    acc = np.zeros(4, np.float64)
    for v in s:  # front to back over the same samples
        e = lut[int(v)].astype(np.float64)
        w = (1 - acc[3]) * e[3]
        acc[:3] += w * e[:3]
        acc[3] += w
    assert np.allclose(o[:3], acc[:3], atol=2e-6) and abs(o[3] - acc[3]) < 2e-6


def _visibility_order(nb, eye_brick):
    ids = [(x, y, z) for z in range(nb) for y in range(nb) for x in range(nb)]
    return sorted(ids, key=lambda b: sum(abs(b[r] - eye_brick[r]) for r in range(3)))


def test_bricked_tiles_equal_unbricked():
    n, B = 48, 16
    gp = configs.gen_params(1)
    vol = ivr.generate_host(gp, n, n, n)
    lut = configs.tf_lut(1)
    st = ivr.state_from_angles(2.2, -0.35, 40, 40)
    v = oracle.view_of(st)
    full, s_full = oracle.render(vol, v, lut, oracle.params(ert=2.0))
    cam = np.zeros(23, np.float32)
    rect = (C.c_int * 4)()
    bmin, bmax = oracle.box(n, n, n)
    oracle.lib.ivro_camera(f(bmin), f(bmax), C.byref(v), f(cam), rect)
    eb = [int(np.floor((cam[r] - bmin[r]) * n / (bmax[r] - bmin[r]) / B)) for r in range(3)]
    nb = (n - 2) // B + 1
    tiles, s_tiles = [], 0
    for b in _visibility_order(nb, eb):
        t, s = oracle.render(vol, v, lut, oracle.params(ert=2.0, brick=B, brick_id=b))
        tiles.append(t)
        s_tiles += s
    comp = oracle.composite_tiles(np.stack(tiles))
    assert s_tiles == s_full  # every grid sample belongs to exactly one brick
    assert np.abs(comp - full).max() < 2e-6


def test_fp64_oracle_close():
    gp = configs.gen_params(1)
    vol = ivr.generate_host(gp, 64, 64, 64)
    lut = configs.tf_lut(1)
    st = configs.state(1, 64, 64)
    v = oracle.view_of(st)
    p = oracle.params()
    i32, n32 = oracle.render(vol, v, lut, p)
    i64, n64 = oracle.render(vol, v, lut, p, f64=True)
    err = np.abs(i32 - i64).max()
    assert err < 2.0 / 255.0, err
    assert abs(n32 - n64) <= 0.001 * n64


def test_fp64_oracle_tiles_close():
    """the FP64 oracle's tile mode keeps the brick's own samples: per-brick
    tiles agree with the fp32 tiles and their samples partition the grid"""
    n, B = 48, 16
    gp = configs.gen_params(1)
    vol = ivr.generate_host(gp, n, n, n)
    lut = configs.tf_lut(1)
    st = ivr.state_from_angles(2.2, -0.35, 40, 40)
    v = oracle.view_of(st)
    _, s_full = oracle.render(vol, v, lut, oracle.params(ert=2.0), f64=True)
    nb = (n - 2) // B + 1
    s_tiles = 0
    for b in [(x, y, z) for z in range(nb) for y in range(nb) for x in range(nb)]:
        t32, _ = oracle.render(vol, v, lut, oracle.params(ert=2.0, brick=B, brick_id=b))
        t64, s = oracle.render(vol, v, lut, oracle.params(ert=2.0, brick=B, brick_id=b), f64=True)
        s_tiles += s
        assert np.abs(t64 - t32).max() < 1e-3, b
    assert abs(s_tiles - s_full) <= 0.001 * s_full


@pytest.mark.parametrize("n,B", [(48, 16), (64, 32), (61, 16)])
def test_one_pass_sort_last_equals_tiles(n, B):
    """the one-pass sort-last oracle (each ray's per-brick runs composited
    from zero with ERT inside the run, combined far to near) equals the
    per-brick tile renders composited by compositetiles, bit for bit and
    with the same sample count, ERT on; with ERT off it equals the
    unbricked render within rounding"""
    gp = configs.gen_params(2)
    vol = ivr.generate_host(gp, n, n, n)
    lut = configs.tf_lut(2)
    lut[:, 3] = np.minimum(1.0, lut[:, 3] * 12.0)  # opaque enough that ERT ends runs inside bricks
    st = ivr.state_from_angles(0.7, 0.45, 56, 48)
    v = oracle.view_of(st)
    cam = np.zeros(23, np.float32)
    rect = (C.c_int * 4)()
    bmin, bmax = oracle.box(n, n, n)
    oracle.lib.ivro_camera(f(bmin), f(bmax), C.byref(v), f(cam), rect)
    eb = [int(np.floor((cam[r] - bmin[r]) * n / (bmax[r] - bmin[r]) / B)) for r in range(3)]
    nb = (n - 2) // B + 1
    tiles, s_tiles = [], 0
    for b in _visibility_order(nb, eb):
        t, s = oracle.render(vol, v, lut, oracle.params(brick=B, brick_id=b))
        tiles.append(t)
        s_tiles += s
    comp = oracle.composite_tiles(np.stack(tiles))
    one, s_one = oracle.render(vol, v, lut, oracle.params(brick=B, sort_last=True))
    assert s_one == s_tiles
    assert np.array_equal(one.view(np.uint32), comp.view(np.uint32))
    full, _ = oracle.render(vol, v, lut, oracle.params(ert=2.0))
    one_off, _ = oracle.render(vol, v, lut, oracle.params(ert=2.0, brick=B, sort_last=True))
    assert np.abs(one_off - full).max() < 2e-6
    # ERT ended some runs early: fewer samples than without it
    _, s_off = oracle.render(vol, v, lut, oracle.params(ert=2.0, brick=B, sort_last=True))
    assert s_one < s_off
