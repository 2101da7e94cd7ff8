"""Colour tables built on the device: the value histogram of a volume
(histogram, mathutil.c:444-470), the five colormap builders
(colormap.c:21-181) and initvolluts' LUT set-up (main.c:1075-1200), against
the golden colormaps of the compiled reference, a numpy restatement of the
histogram and the compiled reference's own initvolluts."""
import ctypes as C
import os

import numpy as np
import pytest

import ivr_amd as ivr
import oracle
from ivr_amd import configs

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def test_device_colormaps_match_reference_goldens(ctx):
    """all five builders at hsize 256 (range 0..255) and 3518 (1468..4985),
    bit for bit against the tables the compiled reference wrote"""
    z = np.load(os.path.join(G, "colormaps.npz"))
    tf = ctx.tf(np.zeros((1, 4), np.float32))
    for key in z.files:
        cid, hsize = (int(t) for t in key[3:].split("_"))
        hmin, hmax = (0.0, 255.0) if hsize == 256 else (1468.0, 4985.0)
        tf.colormap(cid, hsize, hmin, hmax)
        got = tf.download()
        assert np.array_equal(got.view(np.uint32), z[key].view(np.uint32)), key
    tf.close()


def _np_histogram(v):
    """histogram (mathutil.c:444-470) restated in numpy float32"""
    v = v.astype(np.float32).ravel()
    lo, hi = np.float32(v.min()), np.float32(v.max())
    hs = int(np.float32(np.float32(hi - lo) + np.float32(1.0)))
    idx = (v - lo).astype(np.float32).astype(np.int64)
    return lo, hi, hs, np.bincount(idx, minlength=hs).astype(np.int32)


@pytest.mark.parametrize("kind,dtype,dims,brick", [
    ("c1x255", ivr.IVR_F32, (64, 64, 64), 32),
    ("shell", ivr.IVR_F32, (61, 47, 53), 16),
    ("shell", ivr.IVR_F16, (70, 66, 40), 32),
])
def test_histogram_matches_reference_formula(ctx, kind, dtype, dims, brick):
    """min, max, hsize = (int)(max - min + 1) and the counts of (int)(v - min)
    over every voxel exactly once, bricked, fp32 and fp16 storage"""
    nx, ny, nz = dims
    blobs = ivr.generate_host(configs.gen_params(1), nx, ny, nz)
    if kind == "c1x255":
        host = (blobs * np.float32(255.0)).astype(np.float32)
    else:
        host = (np.float32(1468.0) + np.float32(3518.0) * blobs).astype(np.float32)
    vol = ctx.volume(nx, ny, nz, dtype, brick=brick)
    vol.upload(host)
    stored = vol.download()  # fp16 volumes hold the rounded values
    lo, hi, hs, counts = vol.histogram()
    rlo, rhi, rhs, rcounts = _np_histogram(stored)
    print("%s: %g .. %g, %d bins" % (kind, lo, hi, hs))
    assert (np.float32(lo), np.float32(hi), hs) == (rlo, rhi, rhs)
    assert counts.sum() == nx * ny * nz
    assert np.array_equal(counts, rcounts)
    vol.close()


def test_initvolluts_matches_the_compiled_reference(ctx, tmp_path):
    """ivr_volume_luts == the reference's init -> initvolluts on the same raw
    file: every colormap's table and lutoffset; rainbow adds its ramp to the
    table the reference mallocs, so only its entry 0 and opacities are defined
    and compared"""
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("the compiled reference (oracle/_ref) is not built")
    nx, ny, nz = 61, 47, 53
    blobs = ivr.generate_host(configs.gen_params(1), nx, ny, nz)
    host = (np.float32(1468.0) + np.float32(3518.0) * blobs).astype(np.float32)
    path = str(tmp_path / "shell.raw")
    host.tofile(path)
    vol = ctx.volume(nx, ny, nz, ivr.IVR_F32, brick=32)
    vol.upload(host)
    tf = ctx.tf(np.zeros((1, 4), np.float32))
    for cid in range(5):
        off, mx, hs = vol.luts(cid, tf)
        assert ref.ivrref_init(path.encode(), nx, ny, nz, cid, 3, 8) == 0
        want = np.zeros((hs, 4), np.float32)
        ref.ivrref_get_lut(f(want), hs)
        got = tf.download()
        assert off == ref.ivrref_lutoffset() and mx == hs - 1
        if cid == 1:  # rainbow adds to the malloc'ed table: only entry 0 and the opacities are defined
            assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))
            assert np.array_equal(got[:, 3].view(np.uint32), want[:, 3].view(np.uint32))
        else:
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), cid
    tf.close()
    vol.close()
