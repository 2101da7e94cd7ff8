"""Writes the golden vectors of tests/golden/ from the reference renderer
compiled from its own C sources (oracle/_ref/libivrref.so, built by
`make -C oracle ref` where the reference tree exists).

Inputs are synthetic and regenerated from seeds: the C1 Gaussian-blob volume
of ivr_amd's host generator scaled to [0,255] (the reference indexes its LUT
with the raw voxel value, transferfuncs.c:145), whose SHA-256 is stored with
the vectors, orbit cameras in the reference STATE form and LUTs from the
reference's own colormaps or the 5-point builder.

Vectors written (reference function -> file):
  createlocalimage + raycast + CHEM/SEISMIC2  -> compat_images.npz
  getlinearsample (ray.c:87-181)               -> sampler.npz
  getrayintersect (ray.c:276-318)              -> slab.npz
  getsamplesonray (ray.c:239-273)              -> samplesonray.npz
  the six transfer functions                   -> transferfuncs.npz
  the five colormaps (colormap.c)              -> colormaps.npz
  compositetiles (main.c:1203-1218)            -> composite.npz
  getfirstvoxel + getvoxelsonray (ray.c:333-483) -> dda.npz
  getnearestsample (ray.c:182-238)             -> nearest.npz
  createlocalimage with cut planes (STATE newconst, setplaneconst ray.c:44-51),
    a panned eye and a box projected past the viewport -> compat_cuts.npz

usage: python tests/golden/make_golden.py [section ...]   (sections: dda, cuts, nearest; none = all)
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import ivr_amd as ivr  # noqa: E402
from ivr_amd import configs  # noqa: E402
import oracle  # noqa: E402

N = 64
W = H = 96
CAMERAS = [(0.3, 0.2), (1.9, -0.45), (4.0, 0.55)]


def volume():
    gp = configs.gen_params(1)
    gp.value_scale = 255.0
    return ivr.generate_host(gp, N, N, N)


def f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dda(ref):
    """getfirstvoxel / getvoxelsonray over rays given in world and in cell
    coordinates (the reference's walk takes the origin as cell coordinates
    and the entry parameters from the world box), with zero direction
    components and grid-point origins (face ties); at the volume's own
    localdims and at non-cubic ones, which expose the offset's index order."""
    rng = np.random.default_rng(777)
    rays = []
    for k in range(120):
        o = rng.uniform(-1.5, 1.5, 3) if k < 60 else rng.uniform(-4.0, 68.0, 3)
        d = rng.normal(size=3) * rng.uniform(0.2, 3.0)
        rays.append(np.concatenate([o, d]))
    rays = np.array(rays, np.float32)
    for k in range(6):
        rays[k, 3 + k % 3] = 0.0
    for k in range(6, 9):
        rays[k, 3:] = 0.0
        rays[k, 3 + k % 3] = (-1.0) ** k * 0.7
    for k in range(9, 14):
        rays[k, :3] = np.floor(rays[k, :3])
        rays[k, 3:] = np.float32([1.0, -1.0, 1.0]) * (0.5 if k % 2 else 1.0)
    maxdepth = 80
    dims0 = (C.c_int * 3)()
    ref.ivrref_localdims(dims0)
    out = {"rays": rays, "maxdepth": maxdepth}
    rc, first = [], []
    for r in rays:
        v = (C.c_int * 3)()
        rc.append(ref.ivrref_getfirstvoxel(f(r), v))
        first.append(list(v))
    out["rc"] = np.array(rc, np.int32)
    out["first"] = np.array(first, np.int32)
    for tag, dims in (("own", tuple(dims0)), ("odd", (40, 56, 72))):
        ref.ivrref_set_localdims((C.c_int * 3)(*dims))
        walk = np.full((len(rays), maxdepth), -7.0, np.float32)
        for k, r in enumerate(rays):
            buf = np.ascontiguousarray(walk[k])
            ref.ivrref_getvoxelsonray(f(r), maxdepth, f(buf))
            walk[k] = buf
        out["dims_" + tag] = np.array(dims, np.int32)
        out["voxels_" + tag] = walk
    ref.ivrref_set_localdims(dims0)
    box = np.zeros(6, np.float32)
    ref.ivrref_box(f(box))
    out["box"] = box
    np.savez_compressed(os.path.join(HERE, "dda.npz"), **out)


CUT_CASES = [  # (az, el, modelview z translation, x/y pan, newconst)
    (0.7, 0.3, -5.0, (0.0, 0.0), [0.35, 0.8, 10, 0.5, 10, 0.6]),
    (1.1, -0.2, -0.6, (0.0, 0.0), [10] * 6),
    (1.1, -0.2, -5.0, (0.35, -0.25), [10] * 6),
    (2.0, 0.4, -1.2, (-0.3, 0.2), [0.3, 10, 0.2, 10, 0.45, 10]),
]


def cut_states():
    """STATEs of CUT_CASES: the orbit camera with its modelview translation
    replaced and the cut planes set.  buildray takes the eye from the x/y
    translation only (mathutil.c:1129-1137), so a pan moves the rays while
    the z translation moves only the projected box (updatevolbbox,
    mathutil.c:948-1018), here to a rectangle over the whole viewport."""
    out = []
    for az, el, tz, pan, nc in CUT_CASES:
        st = ivr.state_from_angles(az, el, W, H)
        st.alpha = 1.0
        st.mvm[12], st.mvm[13], st.mvm[14] = pan[0], pan[1], tz
        for i, v in enumerate(nc):
            st.newconst[i] = v
        out.append(st)
    return out


def cuts(ref, lut):
    """compat frames (CHEM, the 5-point LUT) with cut planes, a panned eye
    and a projected box larger than the viewport"""
    imgs, states = [], []
    ref.ivrref_set_lut(f(lut), 256, 0, 255)
    ref.ivrref_set_xfer(3)
    for st in cut_states():
        out = np.zeros((H, W, 4), np.float32)
        rect = (C.c_int * 4)()
        ref.ivrref_render(f(np.array(st.mvm, np.float32)), f(np.array(st.pm, np.float32)),
                          f(np.array(st.vv, np.float32)), f(np.array(st.newconst, np.float32)),
                          C.c_float(st.alpha), W, H, f(out), rect)
        imgs.append(out)
        states.append(np.concatenate([st.mvm, st.pm, st.vv, st.newconst, [st.alpha]]).astype(np.float32))
    np.savez_compressed(os.path.join(HERE, "compat_cuts.npz"), images=np.stack(imgs), states=np.stack(states),
                        lut=lut)


def nearest(ref):
    """getnearestsample at points inside the box and on its faces (the
    reference has no lower index clamp, so no point lies below the box)"""
    rng = np.random.default_rng(778)
    pts = rng.uniform(-1.0, 1.0, (200, 3)).astype(np.float32)
    pts[:20, 0] = -1.0
    pts[20:40, 1] = 1.0
    pts[40:60, 2] = 0.999
    ref.ivrref_getnearestsample.restype = C.c_float
    vals = np.array([ref.ivrref_getnearestsample(C.c_float(p[0]), C.c_float(p[1]), C.c_float(p[2]))
                     for p in pts], np.float32)
    np.savez_compressed(os.path.join(HERE, "nearest.npz"), points=pts, values=vals)


def main(sections=()):
    ref = oracle.ref_lib()
    if ref is None:
        sys.exit("oracle/_ref/libivrref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(12345)
    vol = volume()
    sha = hashlib.sha256(vol.tobytes()).hexdigest()
    path = "/tmp/ivr_golden_vol.raw"
    vol.tofile(path)
    ref.ivrref_init(path.encode(), N, N, N, -1, 3, 8)
    if sections:
        if "dda" in sections:
            dda(ref)
        if "cuts" in sections:
            cuts(ref, configs.tf_lut(1))
        if "nearest" in sections:
            nearest(ref)
        print("wrote", ", ".join(sections))
        return
    box = np.zeros(6, np.float32)
    ref.ivrref_box(f(box))
    meta = dict(volume="C1 blobs seed 1001 x255", n=N, sha256=sha, width=W, height=H,
                box=box.tolist(), cameras=CAMERAS)

    # LUTs: the 5-point builder and the reference's medical/jet colormaps
    luts = {"tf5": configs.tf_lut(1)}
    for name, cid in (("medical", 2), ("jet", 4)):
        out = np.zeros((256, 4), np.float32)
        ref.ivrref_colormap(cid, 256, C.c_float(0.0), C.c_float(255.0), f(out))
        luts[name] = out
    cases = [("tf5", 3, 1.0, 255), ("tf5", 3, 0.25, 255), ("medical", 3, 0.25, 255),
             ("jet", 2, 0.25, 0), ("tf5", 2, 1.0, 200)]
    imgs, rects, states, case_meta = [], [], [], []
    for (lname, xfer, alpha, maxidx) in cases:
        for (az, el) in CAMERAS:
            st = ivr.state_from_angles(az, el, W, H)
            st.alpha = alpha
            mvm = np.array(st.mvm, np.float32)
            pm = np.array(st.pm, np.float32)
            vv = np.array(st.vv, np.float32)
            nc = np.array(st.newconst, np.float32)
            lut = luts[lname]
            ref.ivrref_set_lut(f(lut), 256, 0, maxidx)
            out = np.zeros((H, W, 4), np.float32)
            rect = (C.c_int * 4)()
            ref.ivrref_set_xfer(xfer)
            ref.ivrref_render(f(mvm), f(pm), f(vv), f(nc), C.c_float(alpha), W, H, f(out), rect)
            imgs.append(out)
            rects.append(list(rect))
            states.append(np.concatenate([mvm, pm, vv, nc, [alpha]]).astype(np.float32))
            case_meta.append(dict(lut=lname, xfer=xfer, alpha=alpha, maxlutindx=maxidx, az=az, el=el))
    np.savez_compressed(os.path.join(HERE, "compat_images.npz"), images=np.stack(imgs),
                        rects=np.array(rects, np.int32), states=np.stack(states),
                        lut_tf5=luts["tf5"], lut_medical=luts["medical"], lut_jet=luts["jet"])
    meta["compat_cases"] = case_meta

    # sampler: random points inside (and on the faces of) the box
    pts = rng.uniform(-1.0, 1.0, (300, 3)).astype(np.float32)
    pts[:20, 0] = -1.0
    pts[20:40, 1] = 1.0
    ref.ivrref_getlinearsample.restype = C.c_float
    vals = np.array([ref.ivrref_getlinearsample(C.c_float(p[0]), C.c_float(p[1]), C.c_float(p[2]))
                     for p in pts], np.float32)
    np.savez_compressed(os.path.join(HERE, "sampler.npz"), points=pts, values=vals)

    # slab test and samples on ray: rays from orbit eyes towards jittered targets
    rays = []
    for _ in range(300):
        e = rng.normal(size=3)
        e = e / np.linalg.norm(e) * rng.uniform(2.0, 6.0)
        tgt = rng.uniform(-1.5, 1.5, 3)
        d = (tgt - e) * rng.uniform(0.2, 3.0)
        rays.append(np.concatenate([e, d]))
    rays = np.array(rays, np.float32)
    rays[:10, 3] = 0.0  # axis-parallel directions exercise denom == 0
    hits, tins, touts = [], [], []
    for r in rays:
        ti, to = C.c_float(), C.c_float()
        hits.append(ref.ivrref_getrayintersect(f(r), C.byref(ti), C.byref(to)))
        tins.append(ti.value)
        touts.append(to.value)
    np.savez_compressed(os.path.join(HERE, "slab.npz"), rays=rays, hit=np.array(hits, np.int32),
                        tin=np.array(tins, np.float32), tout=np.array(touts, np.float32))
    sor_n, sor_s = [], []
    for r in rays[:40]:
        buf = np.zeros(256, np.float32)
        n = ref.ivrref_getsamplesonray(f(r), 256, f(buf))
        sor_n.append(n)
        sor_s.append(buf)
    np.savez_compressed(os.path.join(HERE, "samplesonray.npz"), rays=rays[:40],
                        n=np.array(sor_n, np.int32), samples=np.stack(sor_s))

    # transfer functions over random sample arrays, LUT offset 3, maxlutindx 150
    lut = luts["tf5"].copy()
    lut[:, 3] = rng.uniform(0.0, 0.6, 256).astype(np.float32)
    ref.ivrref_set_lut(f(lut), 256, 3, 150)
    tf_samples = rng.uniform(3.0, 250.0, (12, 64)).astype(np.float32)
    tf_out = np.zeros((6, 12, 4), np.float32)
    for x in range(6):
        for k in range(12):
            o = np.zeros(4, np.float32)
            s = tf_samples[k].copy()
            ref.ivrref_transferfunc(x, f(s), 64 - k, C.c_float(0.7), f(o))
            tf_out[x, k] = o
    np.savez_compressed(os.path.join(HERE, "transferfuncs.npz"), lut=lut, lutoffset=3, maxlutindx=150,
                        samples=tf_samples, alpha=0.7, out=tf_out)

    # colormaps at the reference harness sizes
    cmaps = {}
    for hsize, hmin, hmax in ((256, 0.0, 255.0), (3518, 1468.0, 4985.0)):
        for cid in range(5):
            out = np.zeros((hsize, 4), np.float32)
            ref.ivrref_colormap(cid, hsize, C.c_float(hmin), C.c_float(hmax), f(out))
            cmaps["map%d_%d" % (cid, hsize)] = out
    np.savez_compressed(os.path.join(HERE, "colormaps.npz"), **cmaps)

    # compositetiles over 5 random premultiplied tiles of 16x16
    a = rng.uniform(0.0, 1.0, (5, 16, 16, 1)).astype(np.float32)
    c = (rng.uniform(0.0, 1.0, (5, 16, 16, 3)) * a).astype(np.float32)
    tiles = np.concatenate([c, a], axis=-1)
    dst = np.zeros((16, 16, 4), np.float32)
    ref.ivrref_compositetiles(f(dst), f(np.ascontiguousarray(tiles)), 5, 16, 16)
    np.savez_compressed(os.path.join(HERE, "composite.npz"), tiles=tiles, out=dst)

    dda(ref)
    cuts(ref, luts["tf5"])
    nearest(ref)

    with open(os.path.join(HERE, "golden_meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print("wrote golden vectors; volume sha256", sha)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
