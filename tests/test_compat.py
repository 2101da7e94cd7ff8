"""The drop-in library libivr_compat.so (include/ivr_compat.h).

oracle/_ref/libivrref_gpu.so is the reference's own main.c, mathutil.c and
utils.c (compiled from its sources) linked against libivr_compat.so in place
of ray.c, transferfuncs.c and colormap.c.  Driven by the same harness as the
golden vectors, it must reproduce them bit for bit: whole frames (through the
reference's createlocalimage calling our raycast per pixel, and through
ivr_compat_createlocalimage as one launch), samplers, slab tests, sample
marching, the voxel walk, transfer functions and colormaps.
"""
import ctypes as C
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

import ivr_amd as ivr
from ivr_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")
META = json.load(open(os.path.join(G, "golden_meta.json")))
COMPAT = os.path.join(ROOT, "ivr_amd", "libivr_compat.so")
REFGPU = os.path.join(ROOT, "oracle", "_ref", "libivrref_gpu.so")


def f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def test_compat_exports_reference_names():
    src = open(os.path.join(ROOT, "include", "ivr_compat.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"^[\w ]+?\b(\w+)\s*\([^;{]*\)\s*;", src, flags=re.M))
    out = subprocess.run(["nm", "-D", "--defined-only", COMPAT], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert {"raycast", "getsamplesonray", "chemistrytransferfunc", "medcolormap"} <= names
    missing = sorted(names - exported)
    assert not missing, missing
    for g in ("dconst", "getsamplefunc", "transferfunc", "colormap"):
        assert g in exported, g


def _load_refgpu():
    if not os.path.exists(REFGPU):
        pytest.skip("oracle/_ref/libivrref_gpu.so not built (needs the reference sources)")
    lib = C.CDLL(REFGPU)
    lib.ivrref_getlinearsample.restype = C.c_float
    lib.ivrref_getnearestsample.restype = C.c_float
    return lib


@pytest.mark.gpu
def test_colormaps_through_dropin():
    """the drop-in builders (one device launch over the caller's table) give
    the reference's colormaps bit for bit"""
    lib = _load_refgpu()
    z = np.load(os.path.join(G, "colormaps.npz"))
    for key in z.files:
        cid, hsize = (int(t) for t in key[3:].split("_"))
        hmin, hmax = (0.0, 255.0) if hsize == 256 else (1468.0, 4985.0)
        out = np.zeros((hsize, 4), np.float32)
        lib.ivrref_colormap(cid, hsize, C.c_float(hmin), C.c_float(hmax), f(out))
        assert np.array_equal(out.view(np.uint32), z[key].view(np.uint32)), key


@pytest.fixture(scope="module")
def refgpu():
    lib = _load_refgpu()
    gp = configs.gen_params(1)
    gp.value_scale = 255.0
    vol = ivr.generate_host(gp, 64, 64, 64)
    assert hashlib.sha256(vol.tobytes()).hexdigest() == META["sha256"]
    path = "/tmp/ivr_compat_vol.raw"
    vol.tofile(path)
    assert lib.ivrref_init(path.encode(), 64, 64, 64, -1, 3, 8) == 0
    return lib


@pytest.mark.gpu
def test_frames_through_dropin(refgpu):
    z = np.load(os.path.join(G, "compat_images.npz"))
    luts = {"tf5": z["lut_tf5"], "medical": z["lut_medical"], "jet": z["lut_jet"]}
    W, H = META["width"], META["height"]
    stats = lambda: [(refgpu.ivr_compat_raycast_stats(C.byref(a), C.byref(b)), a.value, b.value)[1:]  # noqa: E731
                     for a, b in [(C.c_ulonglong(), C.c_ulonglong())]][0]
    for use_gpu_frame in (1, 0):
        assert refgpu.ivrref_frame_on_gpu(use_gpu_frame) == use_gpu_frame
        cases = list(enumerate(META["compat_cases"]))
        l0, r0 = stats()
        for k, case in cases:
            print("drop-in frame", use_gpu_frame, k, case, flush=True)
            st = ivr.state_from_angles(case["az"], case["el"], W, H)
            lut = luts[case["lut"]]
            refgpu.ivrref_set_lut(f(lut), 256, 0, case["maxlutindx"])
            refgpu.ivrref_set_xfer(case["xfer"])
            out = np.zeros((H, W, 4), np.float32)
            rect = (C.c_int * 4)()
            rc = refgpu.ivrref_render(f(np.array(st.mvm, np.float32)), f(np.array(st.pm, np.float32)),
                                      f(np.array(st.vv, np.float32)), f(np.array(st.newconst, np.float32)),
                                      C.c_float(case["alpha"]), W, H, f(out), rect)
            assert rc == 0
            assert np.array_equal(out.view(np.uint32), z["images"][k].view(np.uint32)), (use_gpu_frame, case)
        l1, r1 = stats()
        if not use_gpu_frame:
            # the reference's pixel loop called raycast from its OpenMP threads; concurrent calls
            # shared launches
            print("per-pixel raycast: %d rays in %d launches" % (r1 - r0, l1 - l0), flush=True)
            assert r1 - r0 > 0 and (os.cpu_count() == 1 or l1 - l0 < r1 - r0)
    refgpu.ivrref_frame_on_gpu(0)


@pytest.mark.gpu
def test_per_ray_functions_through_dropin(refgpu):
    z = np.load(os.path.join(G, "sampler.npz"))
    vals = np.array([refgpu.ivrref_getlinearsample(C.c_float(p[0]), C.c_float(p[1]), C.c_float(p[2]))
                     for p in z["points"][:60]], np.float32)
    assert np.array_equal(vals.view(np.uint32), z["values"][:60].view(np.uint32))
    z = np.load(os.path.join(G, "nearest.npz"))
    vals = np.array([refgpu.ivrref_getnearestsample(C.c_float(p[0]), C.c_float(p[1]), C.c_float(p[2]))
                     for p in z["points"][:60]], np.float32)
    assert np.array_equal(vals.view(np.uint32), z["values"][:60].view(np.uint32))
    z = np.load(os.path.join(G, "slab.npz"))
    for r, hit, tin, tout in list(zip(z["rays"], z["hit"], z["tin"], z["tout"]))[:60]:
        ti, to = C.c_float(), C.c_float()
        assert refgpu.ivrref_getrayintersect(f(np.ascontiguousarray(r)), C.byref(ti), C.byref(to)) == hit
        assert np.float32(ti.value).tobytes() == np.float32(tin).tobytes()
        assert np.float32(to.value).tobytes() == np.float32(tout).tobytes()
    z = np.load(os.path.join(G, "samplesonray.npz"))
    for r, n, s in list(zip(z["rays"], z["n"], z["samples"]))[:20]:
        buf = np.zeros(256, np.float32)
        got = refgpu.ivrref_getsamplesonray(f(np.ascontiguousarray(r)), 256, f(buf))
        assert got == n
        if n > 0:
            assert np.array_equal(buf[:n].view(np.uint32), s[:n].view(np.uint32))
    z = np.load(os.path.join(G, "transferfuncs.npz"))
    refgpu.ivrref_set_lut(f(np.ascontiguousarray(z["lut"])), 256, int(z["lutoffset"]), int(z["maxlutindx"]))
    for x in range(6):
        for k in range(0, 12, 3):
            o = np.zeros(4, np.float32)
            s = np.ascontiguousarray(z["samples"][k])
            refgpu.ivrref_transferfunc(x, f(s), 64 - k, C.c_float(float(z["alpha"])), f(o))
            assert np.array_equal(o.view(np.uint32), z["out"][x, k].view(np.uint32)), (x, k)


@pytest.mark.gpu
def test_voxel_walk_through_dropin(refgpu):
    """getfirstvoxel / getvoxelsonray of the drop-in (one GPU launch per
    call) against the compiled reference's vectors, at the volume's own
    localdims and at non-cubic ones"""
    z = np.load(os.path.join(G, "dda.npz"))
    md = int(z["maxdepth"])
    dims0 = (C.c_int * 3)()
    refgpu.ivrref_localdims(dims0)
    assert list(dims0) == list(z["dims_own"])
    for k, r in enumerate(z["rays"]):
        v = (C.c_int * 3)()
        assert refgpu.ivrref_getfirstvoxel(f(np.ascontiguousarray(r)), v) == z["rc"][k]
        if z["rc"][k] == 0:
            assert list(v) == list(z["first"][k]), k
    zero = np.float32([0.3, 0.2, 0.1, 0.0, 0.0, 0.0])  # all-zero direction: a miss, nothing written
    v = (C.c_int * 3)(5, 6, 7)
    assert refgpu.ivrref_getfirstvoxel(f(zero), v) == -1 and list(v) == [5, 6, 7]
    buf = np.full(4, -7.0, np.float32)
    refgpu.ivrref_getvoxelsonray(f(zero), 4, f(buf))
    assert np.all(buf == -7.0)
    try:
        for tag in ("own", "odd"):
            refgpu.ivrref_set_localdims((C.c_int * 3)(*[int(x) for x in z["dims_" + tag]]))
            print("voxel walk", tag, flush=True)
            for k, r in enumerate(z["rays"]):
                buf = np.full(md, -7.0, np.float32)
                refgpu.ivrref_getvoxelsonray(f(np.ascontiguousarray(r)), md, f(buf))
                assert np.array_equal(buf.view(np.uint32), z["voxels_" + tag][k].view(np.uint32)), (tag, k)
    finally:
        refgpu.ivrref_set_localdims(dims0)


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


@pytest.mark.gpu
def test_cut_frames_through_dropin(refgpu):
    """frames with cut planes, a panned eye and a box projected past the
    viewport through the drop-in: the whole-frame launch for every case, the
    per-pixel raycast path for the first"""
    z = np.load(os.path.join(G, "compat_cuts.npz"))
    W, H = META["width"], META["height"]
    lut = np.ascontiguousarray(z["lut"])
    for use_gpu_frame in (1, 0):
        assert refgpu.ivrref_frame_on_gpu(use_gpu_frame) == use_gpu_frame
        rows = list(enumerate(z["states"]))
        for k, row in (rows if use_gpu_frame else rows[:1]):
            print("drop-in cut/pan frame", use_gpu_frame, k, flush=True)
            st = _state_of(row, W, H)
            refgpu.ivrref_set_lut(f(lut), 256, 0, 255)
            refgpu.ivrref_set_xfer(3)
            out = np.zeros((H, W, 4), np.float32)
            rect = (C.c_int * 4)()
            rc = refgpu.ivrref_render(f(np.array(st.mvm, np.float32)), f(np.array(st.pm, np.float32)),
                                      f(np.array(st.vv, np.float32)), f(np.array(st.newconst, np.float32)),
                                      C.c_float(st.alpha), W, H, f(out), rect)
            assert rc == 0
            assert np.array_equal(out.view(np.uint32), z["images"][k].view(np.uint32)), (use_gpu_frame, k)
    refgpu.ivrref_frame_on_gpu(0)


@pytest.mark.gpu
def test_compositetiles_through_dropin(refgpu):
    """ivr_compat_compositetiles (compositetiles, main.c:1203-1218, on the GPU)
    gives the compiled reference's output bit for bit: on the golden tiles (5
    premultiplied 16x16 tiles over a zero buffer) and on ragged 37x23 tiles
    over a non-zero buffer against the compiled reference run on the same
    inputs"""
    import oracle
    z = np.load(os.path.join(G, "composite.npz"))
    fn = refgpu.ivr_compat_compositetiles  # libivr_compat.so, loaded by the relinked host
    dst = np.zeros((16, 16, 4), np.float32)
    assert fn(f(dst), f(np.ascontiguousarray(z["tiles"])), 5, 16, 16) == 0
    assert np.array_equal(dst.view(np.uint32), z["out"].view(np.uint32))
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libivrref.so not built")
    rng = np.random.default_rng(12)
    a = rng.uniform(0.0, 1.0, (3, 23, 37, 1)).astype(np.float32)
    tiles = np.ascontiguousarray(np.concatenate([(rng.uniform(0.0, 1.0, (3, 23, 37, 3)) * a).astype(np.float32), a], -1))
    base = np.ascontiguousarray(rng.uniform(0.0, 0.5, (23, 37, 4)).astype(np.float32))
    want = base.copy()
    ref.ivrref_compositetiles(f(want), f(tiles.copy()), 3, 37, 23)
    got = base.copy()
    assert fn(f(got), f(tiles), 3, 37, 23) == 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
