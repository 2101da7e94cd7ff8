"""ivr_render_noalloc: the render entry point that allocates nothing (its
frame counters and sort-last planes live in the library's static scratch
array) gives the images of ivr_render, on consecutive calls that reuse the
array with different and smaller inputs, and refuses frames that do not fit
the array with an error."""
import ctypes as C

import numpy as np
import pytest

import ivr_amd as ivr
from ivr_amd import configs

pytestmark = pytest.mark.gpu

_hip = C.CDLL("libamdhip64.so")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
_hip.hipDeviceSynchronize.argtypes = []


class DevBuf:
    """device memory the caller owns"""

    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.p), nbytes) == 0
        assert _hip.hipMemset(self.p, 0xA5, nbytes) == 0  # stale contents the render must overwrite
        self.n = nbytes

    def get(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.n
        assert _hip.hipMemcpy(C.c_void_p(out.ctypes.data), self.p, self.n, 2) == 0  # device to host
        return out

    def close(self):
        _hip.hipFree(self.p)


def _library_frame(ctx, vol, tf, desc):
    W, H = desc.state.winwidth, desc.state.winheight
    fr = ctx.frame(W, H, keep_rgba32=True)
    ctx.render(vol, tf, desc, fr)
    img8, img32 = fr.download(rgba32=True)
    fr.close()
    return img8, img32


@pytest.mark.parametrize("composite", [ivr.FUSED, ivr.SORT_LAST])
def test_noalloc_equals_render(ctx, composite):
    cases = [  # a larger volume and frame, then a smaller one: the array keeps the first call's contents
        (configs.gen_params(2), 160, 32, 2, 232, 200, True),
        (configs.gen_params(1), 64, 32, 1, 120, 96, False),
    ]
    vols = []
    for gp, n, B, c, W, H, skip in cases:
        vol = ctx.volume(n, n, n, ivr.IVR_F32, brick=B)
        vol.generate(gp)
        vols.append(vol)
    for (gp, n, B, c, W, H, skip), vol in zip(cases, vols):
        print("noalloc case: composite %d, %d^3 bricks %d, %dx%d, skip %s" % (composite, n, B, W, H, skip),
              flush=True)
        tf = ctx.tf(configs.tf_lut(c))
        desc = ivr.config_desc(configs.state(c, W, H), skip=skip, composite=composite)
        ref8, ref32 = _library_frame(ctx, vol, tf, desc)
        before = ctx.device_bytes()
        b8, b32 = DevBuf(W * H * 4), DevBuf(W * H * 16)
        for rep in range(2):
            assert ctx.render_noalloc(vol, tf, desc, b8.p.value, b32.p.value) == 0, ivr.lib.ivr_last_error()
            ctx.sync()
            got8 = b8.get(np.uint8, (H, W, 4))
            got32 = b32.get(np.float32, (H, W, 4))
            assert np.array_equal(got8, ref8), rep
            assert np.array_equal(got32.view(np.uint32), ref32.view(np.uint32)), rep
        # the RGBA32F image is optional
        assert ctx.render_noalloc(vol, tf, desc, b8.p.value) == 0
        ctx.sync()
        assert np.array_equal(b8.get(np.uint8, (H, W, 4)), ref8)
        assert ctx.device_bytes() == before  # nothing allocated through the library
        b8.close()
        b32.close()
        tf.close()
    for vol in vols:
        vol.close()


def test_noalloc_rejects_what_does_not_fit(ctx):
    vol = ctx.volume(96, 96, 96, ivr.IVR_F32, brick=32)  # 3x3x3 bricks: 7 run planes
    vol.generate(configs.gen_params(1))
    tf = ctx.tf(configs.tf_lut(1))
    b8 = DevBuf(2048 * 1024 * 4)  # room for the image, should the check ever let it through
    big = ivr.config_desc(configs.state(1, 2048, 1024), composite=ivr.SORT_LAST)  # 7 x 32 MiB of planes
    assert ctx.render_noalloc(vol, tf, big, b8.p.value) == -1
    assert b"scratch" in ivr.lib.ivr_last_error()
    tiles = ivr.config_desc(configs.state(1, 64, 64), composite=ivr.SORT_LAST_TILES)
    assert ctx.render_noalloc(vol, tf, tiles, b8.p.value) == -4
    b8.close()
    tf.close()
    vol.close()
