"""Python mirror of the libivr.so C-ABI (include/ivr_gpu.h).

The reference drives its renderer from C through globals and a per-frame
STATE (connection.h:9-18); this module exposes the same render path to
Python for tests, the benchmark and the steering server: a Context owns the
HIP stream, Volume/TF/Frame wrap the device objects, and Context.render
enqueues one frame.  Every call goes through the compiled library; there is
no Python or CPU fallback, and importing this module fails loudly when
libivr.so is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libivr.so")

IVR_F32, IVR_F16 = 0, 1
MODE_COMPAT, MODE_CONFIG = 0, 1
FUSED, SORT_LAST, SORT_LAST_TILES = 0, 1, 2
GEN_BLOBS, GEN_NOISE_BLOBS = 0, 1

# reference transfer function ids (transferfuncs.h:8-13)
SEISMIC, MEDXFER, SEISMIC2, CHEM, BLEND, RGBBLEND = 0, 1, 2, 3, 4, 5


class IvrError(RuntimeError):
    pass


class State(C.Structure):
    """STATE (connection.h:9-18), 184 bytes."""
    _fields_ = [("finish", C.c_int), ("winwidth", C.c_int), ("winheight", C.c_int),
                ("alpha", C.c_float), ("newconst", C.c_float * 6), ("vv", C.c_float * 4),
                ("mvm", C.c_float * 16), ("pm", C.c_float * 16)]


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("nblobs", C.c_int), ("octaves", C.c_int),
                ("lacunarity", C.c_float), ("gain", C.c_float), ("base_freq", C.c_float),
                ("empty_frac", C.c_float), ("value_scale", C.c_float), ("seed", C.c_uint64)]


class FrameDesc(C.Structure):
    _fields_ = [("state", State), ("mode", C.c_int), ("xfer", C.c_int), ("step_vox", C.c_float),
                ("ert", C.c_float), ("lutscale", C.c_float), ("lutoffset", C.c_int),
                ("maxlutindx", C.c_int), ("skip", C.c_int), ("composite", C.c_int)]


class FrameStats(C.Structure):
    _fields_ = [("samples", C.c_ulonglong), ("visited", C.c_ulonglong),
                ("render_ms", C.c_float), ("composite_ms", C.c_float), ("partials", C.c_ulonglong),
                ("evaluations", C.c_ulonglong)]


class SurrogateParams(C.Structure):
    _fields_ = [("log2_table", C.c_int), ("base_res", C.c_int), ("growth", C.c_float), ("lr", C.c_float),
                ("lr_mlp", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("init_scale", C.c_float), ("seed", C.c_uint64), ("log2_group", C.c_int)]


SG_NPARAM = 64 * 32 + 64 + 64 * 64 + 64 + 64 + 1


class Request(C.Structure):
    _fields_ = [("id", C.c_ulonglong), ("volume", C.c_int), ("desc", FrameDesc), ("lut", C.c_void_p),
                ("lutsize", C.c_int), ("rgba8", C.c_void_p), ("t_submit", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("id", C.c_ulonglong), ("status", C.c_int), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("samples", C.c_ulonglong)]


class BatcherStats(C.Structure):
    _fields_ = [("submitted", C.c_ulonglong), ("admitted", C.c_ulonglong), ("completed", C.c_ulonglong),
                ("rejected", C.c_ulonglong), ("batches", C.c_ulonglong), ("bytes_in_use", C.c_ulonglong),
                ("budget", C.c_ulonglong), ("pending", C.c_int), ("in_flight", C.c_int)]

assert C.sizeof(State) == 184
assert C.sizeof(GenParams) == 40
assert C.sizeof(FrameDesc) == 220


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libivr.so not built at %s: run __graft_entry__.build() "
                          "(python ivr_amd/build.py)" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "ivr_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
        "ivr_ctx_destroy": (C.c_int, [P]),
        "ivr_ctx_sync": (C.c_int, [P]),
        "ivr_ctx_record": (C.c_int, [P, C.c_int]),
        "ivr_ctx_elapsed": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_float)]),
        "ivr_ctx_device_bytes": (C.c_ulonglong, [P]),
        "ivr_last_error": (C.c_char_p, []),
        "ivr_volume_create": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(P)]),
        "ivr_volume_upload": (C.c_int, [P, P, P]),
        "ivr_volume_generate": (C.c_int, [P, P, C.POINTER(GenParams)]),
        "ivr_volume_download": (C.c_int, [P, P, P]),
        "ivr_volume_gradient": (C.c_int, [P, P, P, C.c_int]),
        "ivr_volume_destroy": (C.c_int, [P, P]),
        "ivr_volume_info": (C.c_int, [P, P, P, P, P, P]),
        "ivr_tf_create": (C.c_int, [P, P, C.c_int, C.POINTER(P)]),
        "ivr_tf_update": (C.c_int, [P, P, P, C.c_int]),
        "ivr_tf_destroy": (C.c_int, [P, P]),
        "ivr_tf_colormap": (C.c_int, [P, P, C.c_int, C.c_int, C.c_float, C.c_float]),
        "ivr_tf_download": (C.c_int, [P, P, P, C.c_int]),
        "ivr_volume_histogram": (C.c_int, [P, P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int),
                                           P, C.c_int]),
        "ivr_volume_luts": (C.c_int, [P, P, C.c_int, P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
        "ivr_frame_create": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(P)]),
        "ivr_frame_destroy": (C.c_int, [P, P]),
        "ivr_frame_download": (C.c_int, [P, P, P, P]),
        "ivr_frame_stats_get": (C.c_int, [P, P, C.POINTER(FrameStats)]),
        "ivr_render": (C.c_int, [P, P, P, C.POINTER(FrameDesc), P]),
        "ivr_render_noalloc": (C.c_int, [P, P, P, C.POINTER(FrameDesc), P, P]),
        "ivr_desc_config": (None, [C.POINTER(FrameDesc)]),
        "ivr_state_orbit": (None, [C.c_uint64, C.c_int, C.c_int, C.POINTER(State),
                                   C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "ivr_state_from_angles": (None, [C.c_float, C.c_float, C.c_int, C.c_int, C.POINTER(State)]),
        "ivr_tf_build": (None, [C.c_uint64, C.c_float, C.c_float, C.c_uint64, P]),
        "ivr_generate_host": (C.c_int, [C.POINTER(GenParams), C.c_int, C.c_int, C.c_int, P]),
        "ivr_rng_u32": (C.c_uint32, [C.POINTER(C.c_uint64)]),
        "ivr_surrogate_defaults": (None, [C.POINTER(SurrogateParams)]),
        "ivr_surrogate_create": (C.c_int, [P, P, C.POINTER(SurrogateParams), C.POINTER(P)]),
        "ivr_surrogate_destroy": (C.c_int, [P, P]),
        "ivr_surrogate_train": (C.c_int, [P, P, P, C.c_int, C.c_longlong, C.POINTER(C.c_float)]),
        "ivr_surrogate_grad": (C.c_int, [P, P, C.c_int, P, P, P, P, C.POINTER(C.c_float), P, P, C.c_int]),
        "ivr_surrogate_adam": (C.c_int, [P, P]),
        "ivr_surrogate_get": (C.c_int, [P, P, P, P, C.c_int]),
        "ivr_surrogate_set": (C.c_int, [P, P, P, P, C.c_int]),
        "ivr_surrogate_info": (C.c_int, [P, P, P, P, P]),
        "ivr_surrogate_render": (C.c_int, [P, P, P, P, C.POINTER(FrameDesc), P]),
        "ivr_surrogate_evaluate": (C.c_int, [P, P, P, P, C.POINTER(FrameDesc), C.c_float, P, P]),
        "ivr_surrogate_set_fallback": (C.c_int, [P, P, P]),
        "ivr_surrogate_eval_points": (C.c_int, [P, P, C.c_int, P, P, P]),
        "ivr_surrogate_save": (C.c_int, [P, P, C.c_char_p]),
        "ivr_surrogate_load": (C.c_int, [P, P, C.c_char_p]),
        "ivr_surrogate_freeze": (C.c_int, [P, P]),
        "ivr_batcher_create": (C.c_int, [P, C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.POINTER(P)]),
        "ivr_batcher_destroy": (C.c_int, [P]),
        "ivr_batcher_add_volume": (C.c_int, [P, P]),
        "ivr_batcher_add_surrogate": (C.c_int, [P, P, P]),
        "ivr_batcher_submit": (C.c_int, [P, C.POINTER(Request)]),
        "ivr_batcher_progress": (C.c_int, [P, C.c_int]),
        "ivr_batcher_poll": (C.c_int, [P, C.POINTER(Completion), C.c_int]),
        "ivr_batcher_stats_get": (C.c_int, [P, C.POINTER(BatcherStats)]),
        "ivr_now": (C.c_double, []),
        "ivr_state_pack": (None, [C.POINTER(State), P]),
        "ivr_state_unpack": (None, [P, C.POINTER(State)]),
        "ivr_state_swap_floats": (None, [C.POINTER(State)]),
        "ivr_tiff_encode": (C.c_longlong, [P, C.c_int, C.c_int, C.c_int, P, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def check(rc):
    if rc != 0:
        raise IvrError("libivr error %d: %s" % (rc, lib.ivr_last_error().decode()))
    return rc


# ---------------------------------------------------------------- host helpers

def orbit_state(seed, width, height):
    st = State()
    az, el = C.c_float(), C.c_float()
    lib.ivr_state_orbit(seed, width, height, C.byref(st), C.byref(az), C.byref(el))
    return st


def state_from_angles(az, el, width, height):
    st = State()
    lib.ivr_state_from_angles(az, el, width, height, C.byref(st))
    return st


def tf_build(seed, alpha_max=0.05, jitter=0.0, jitter_seed=0):
    lut = np.zeros((256, 4), np.float32)
    lib.ivr_tf_build(seed, alpha_max, jitter, jitter_seed, _ptr(lut))
    return lut


def gen_params(kind, seed, nblobs, octaves=4, lacunarity=2.0, gain=0.5, base_freq=4.0,
               empty_frac=0.0, value_scale=1.0):
    return GenParams(kind=kind, nblobs=nblobs, octaves=octaves, lacunarity=lacunarity, gain=gain,
                     base_freq=base_freq, empty_frac=empty_frac, value_scale=value_scale, seed=seed)


def generate_host(gp, nx, ny, nz):
    out = np.empty((nz, ny, nx), np.float32)
    check(lib.ivr_generate_host(C.byref(gp), nx, ny, nz, _ptr(out)))
    return out


def config_desc(state, skip=False, composite=FUSED, ert=0.99, step_vox=0.5, lutscale=255.0):
    d = FrameDesc()
    lib.ivr_desc_config(C.byref(d))
    d.state = state
    d.skip = int(skip)
    d.composite = composite
    d.ert = ert
    d.step_vox = step_vox
    d.lutscale = lutscale
    return d


def compat_desc(state, xfer=CHEM, lutoffset=0, maxlutindx=0):
    d = FrameDesc()
    lib.ivr_desc_config(C.byref(d))
    d.state = state
    d.mode = MODE_COMPAT
    d.xfer = xfer
    d.lutscale = 1.0
    d.lutoffset = lutoffset
    d.maxlutindx = maxlutindx
    d.ert = 2.0
    return d


# ---------------------------------------------------------------- device objects

class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib.ivr_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib.ivr_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        check(lib.ivr_ctx_sync(self.h))

    def record(self, slot):
        check(lib.ivr_ctx_record(self.h, slot))

    def elapsed(self, s0, s1):
        ms = C.c_float()
        check(lib.ivr_ctx_elapsed(self.h, s0, s1, C.byref(ms)))
        return ms.value

    def device_bytes(self):
        return lib.ivr_ctx_device_bytes(self.h)

    def volume(self, nx, ny, nz, dtype=IVR_F32, brick=256):
        return Volume(self, nx, ny, nz, dtype, brick)

    def tf(self, lut):
        return TF(self, lut)

    def frame(self, width, height, keep_rgba32=False):
        return Frame(self, width, height, keep_rgba32)

    def render(self, vol, tf, desc, frame):
        check(lib.ivr_render(self.h, vol.h, tf.h, C.byref(desc), frame.h))

    def render_noalloc(self, vol, tf, desc, d_rgba8, d_rgba32=None):
        """ivr_render_noalloc into caller device buffers (addresses as ints);
        returns the status code (0 or a negative IVR_ERR_*) without raising"""
        return lib.ivr_render_noalloc(self.h, vol.h, tf.h, C.byref(desc), C.c_void_p(d_rgba8),
                                      None if d_rgba32 is None else C.c_void_p(d_rgba32))

    def surrogate(self, vol, **kw):
        return Surrogate(self, vol, **kw)

    def batcher(self, budget_bytes, max_slots, max_width, max_height):
        return Batcher(self, budget_bytes, max_slots, max_width, max_height)


class Volume:
    def __init__(self, ctx, nx, ny, nz, dtype, brick):
        self.ctx = ctx
        self.shape = (nz, ny, nx)
        h = C.c_void_p()
        check(lib.ivr_volume_create(ctx.h, nx, ny, nz, dtype, brick, C.byref(h)))
        self.h = h

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        assert a.shape == self.shape, (a.shape, self.shape)
        check(lib.ivr_volume_upload(self.ctx.h, self.h, _ptr(a)))

    def generate(self, gp):
        check(lib.ivr_volume_generate(self.ctx.h, self.h, C.byref(gp)))

    def download(self):
        out = np.empty(self.shape, np.float32)
        check(lib.ivr_volume_download(self.ctx.h, self.h, _ptr(out)))
        return out

    def gradient(self, device_ptr=None):
        """central-difference gradient field (nz, ny, nx, 4): d/dx, d/dy, d/dz,
        magnitude; into device memory at device_ptr (stream-ordered) when given"""
        if device_ptr is not None:
            check(lib.ivr_volume_gradient(self.ctx.h, self.h, C.c_void_p(device_ptr), 1))
            return None
        out = np.empty(self.shape + (4,), np.float32)
        check(lib.ivr_volume_gradient(self.ctx.h, self.h, _ptr(out), 0))
        return out

    def histogram(self, bins=True):
        """(vmin, vmax, hsize, counts or None): the reference's histogram of the values"""
        lo, hi, hs = C.c_float(), C.c_float(), C.c_int()
        check(lib.ivr_volume_histogram(self.ctx.h, self.h, C.byref(lo), C.byref(hi), C.byref(hs), None, 0))
        counts = None
        if bins:
            counts = np.zeros(hs.value, np.int32)
            check(lib.ivr_volume_histogram(self.ctx.h, self.h, C.byref(lo), C.byref(hi), C.byref(hs),
                                           _ptr(counts), hs.value))
        return lo.value, hi.value, hs.value, counts

    def luts(self, colormap_id, tf):
        """initvolluts: colormap over the value range into tf; returns (lutoffset, maxlutindx, hsize)"""
        off, mx, hs = C.c_int(), C.c_int(), C.c_int()
        check(lib.ivr_volume_luts(self.ctx.h, self.h, colormap_id, tf.h, C.byref(off), C.byref(mx), C.byref(hs)))
        tf.n = hs.value
        return off.value, mx.value, hs.value

    def info(self):
        dims = (C.c_int * 3)()
        nb = (C.c_int * 3)()
        brick, dtype = C.c_int(), C.c_int()
        nbytes = C.c_ulonglong()
        check(lib.ivr_volume_info(self.h, dims, nb, C.byref(brick), C.byref(dtype), C.byref(nbytes)))
        return dict(dims=tuple(dims), nbricks=tuple(nb), brick=brick.value, dtype=dtype.value,
                    bytes=nbytes.value)

    def close(self):
        if self.h:
            lib.ivr_volume_destroy(self.ctx.h, self.h)
            self.h = None


class TF:
    def __init__(self, ctx, lut):
        self.ctx = ctx
        a = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1, 4)
        h = C.c_void_p()
        check(lib.ivr_tf_create(ctx.h, _ptr(a), a.shape[0], C.byref(h)))
        self.h = h
        self.n = a.shape[0]

    def update(self, lut):
        a = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1, 4)
        check(lib.ivr_tf_update(self.ctx.h, self.h, _ptr(a), a.shape[0]))
        self.n = a.shape[0]

    def colormap(self, cid, hsize, hmin, hmax):
        """the reference colormap builder `cid` of hsize entries, on the device"""
        check(lib.ivr_tf_colormap(self.ctx.h, self.h, cid, hsize, hmin, hmax))
        self.n = hsize

    def download(self, n=None):
        n = self.n if n is None else n
        out = np.zeros((n, 4), np.float32)
        check(lib.ivr_tf_download(self.ctx.h, self.h, _ptr(out), n))
        return out

    def close(self):
        if self.h:
            lib.ivr_tf_destroy(self.ctx.h, self.h)
            self.h = None


class Frame:
    def __init__(self, ctx, width, height, keep_rgba32):
        self.ctx = ctx
        self.width, self.height = width, height
        self.keep32 = keep_rgba32
        h = C.c_void_p()
        check(lib.ivr_frame_create(ctx.h, width, height, int(keep_rgba32), C.byref(h)))
        self.h = h

    def download(self, rgba32=False, out=None):
        """RGBA8 image (row 0 = bottom) into `out` (a C-contiguous uint8 array of the
        frame's shape, reused across frames) or a new array; with rgba32 also RGBA32F"""
        shape = (self.height, self.width, 4)
        if out is not None and (out.shape != shape or out.dtype != np.uint8
                                or not out.flags["C_CONTIGUOUS"] or not out.flags["WRITEABLE"]):
            raise ValueError("out must be a writeable C-contiguous uint8 array of shape %s" % (shape,))
        img8 = np.empty(shape, np.uint8) if out is None else out
        img32 = np.empty((self.height, self.width, 4), np.float32) if rgba32 else None
        check(lib.ivr_frame_download(self.ctx.h, self.h, _ptr(img8),
                                     _ptr(img32) if rgba32 else None))
        return (img8, img32) if rgba32 else img8

    def stats(self):
        st = FrameStats()
        check(lib.ivr_frame_stats_get(self.ctx.h, self.h, C.byref(st)))
        return dict(samples=st.samples, visited=st.visited, render_ms=st.render_ms,
                    composite_ms=st.composite_ms, partials=st.partials, evaluations=st.evaluations)

    def close(self):
        if self.h:
            lib.ivr_frame_destroy(self.ctx.h, self.h)
            self.h = None


def surrogate_defaults(**kw):
    sp = SurrogateParams()
    lib.ivr_surrogate_defaults(C.byref(sp))
    for k, v in kw.items():
        setattr(sp, k, v)
    return sp


class Surrogate:
    """per-brick hash grids + shared MLP trained on a plain volume"""

    def __init__(self, ctx, vol, **kw):
        self.ctx = ctx
        self.sp = surrogate_defaults(**kw)
        h = C.c_void_p()
        check(lib.ivr_surrogate_create(ctx.h, vol.h, C.byref(self.sp), C.byref(h)))
        self.h = h
        nb, T, step = C.c_int(), C.c_ulonglong(), C.c_int()
        res = (C.c_int * 16)()
        check(lib.ivr_surrogate_info(h, C.byref(nb), C.byref(T), C.byref(step), res))
        self.nbricks, self.T, self.res = nb.value, T.value, list(res)

    def train(self, vol, steps, samples):
        loss = C.c_float()
        check(lib.ivr_surrogate_train(self.ctx.h, self.h, vol.h, steps, samples, C.byref(loss)))
        return loss.value

    def grad(self, xyz, bricks, targets, table_brick=0):
        xyz = np.ascontiguousarray(xyz, np.float32)
        bricks = np.ascontiguousarray(bricks, np.int32)
        targets = np.ascontiguousarray(targets, np.float32)
        n = len(targets)
        y = np.zeros(n, np.float32)
        g = np.zeros(SG_NPARAM, np.float32)
        gt = np.zeros((16, self.T, 2), np.float32)
        loss = C.c_float()
        check(lib.ivr_surrogate_grad(self.ctx.h, self.h, n, _ptr(xyz), _ptr(bricks), _ptr(targets), _ptr(y),
                                     C.byref(loss), _ptr(g), _ptr(gt), table_brick))
        return y, loss.value, g, gt

    def adam(self):
        check(lib.ivr_surrogate_adam(self.ctx.h, self.h))

    def get(self, brick=0, table=True):
        mlp = np.zeros(SG_NPARAM, np.float32)
        tb = np.zeros((16, self.T, 2), np.float32) if table else None
        check(lib.ivr_surrogate_get(self.ctx.h, self.h, _ptr(mlp), _ptr(tb) if table else None, brick))
        return mlp, tb

    def set(self, mlp=None, table=None, brick=0):
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        t = None if table is None else np.ascontiguousarray(table, np.float32)
        check(lib.ivr_surrogate_set(self.ctx.h, self.h, None if m is None else _ptr(m),
                                    None if t is None else _ptr(t), brick))

    def render(self, vol, tf, desc, frame):
        check(lib.ivr_surrogate_render(self.ctx.h, self.h, None if vol is None else vol.h, tf.h,
                                       C.byref(desc), frame.h))

    def evaluate(self, vol, tf, desc, threshold_db=40.0):
        psnr = np.zeros(self.nbricks, np.float32)
        nf = C.c_int()
        check(lib.ivr_surrogate_evaluate(self.ctx.h, self.h, vol.h, tf.h, C.byref(desc), threshold_db,
                                         _ptr(psnr), C.byref(nf)))
        return psnr, nf.value

    def eval_points(self, xyz, bricks):
        """the renderer's network value at local points xyz (n, 3) of bricks (n,)"""
        xyz = np.ascontiguousarray(xyz, np.float32)
        bricks = np.ascontiguousarray(bricks, np.int32)
        y = np.zeros(len(bricks), np.float32)
        check(lib.ivr_surrogate_eval_points(self.ctx.h, self.h, len(bricks), _ptr(xyz), _ptr(bricks), _ptr(y)))
        return y

    def set_fallback(self, flags):
        f = np.ascontiguousarray(flags, np.uint8)
        check(lib.ivr_surrogate_set_fallback(self.ctx.h, self.h, _ptr(f)))

    def save(self, path):
        check(lib.ivr_surrogate_save(self.ctx.h, self.h, path.encode()))

    def load(self, path):
        check(lib.ivr_surrogate_load(self.ctx.h, self.h, path.encode()))

    def freeze(self):
        """release the training state: only the half-precision tables, the MLP
        and the macrocell ranges stay resident"""
        check(lib.ivr_surrogate_freeze(self.ctx.h, self.h))

    def close(self):
        if self.h:
            lib.ivr_surrogate_destroy(self.ctx.h, self.h)
            self.h = None


def psnr(a, b):
    """PSNR in dB of two RGBA8 images"""
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.mean(d * d))
    return 99.0 if mse == 0 else 10.0 * np.log10(255.0 ** 2 / mse)


def ssim(a, b, win=7):
    """mean SSIM (Wang et al. 2004) of the luminance of two RGBA8 images,
    uniform window, C1=(0.01*255)^2, C2=(0.03*255)^2"""
    def lum(x):
        x = x.astype(np.float64)
        return 0.299 * x[..., 0] + 0.587 * x[..., 1] + 0.114 * x[..., 2]
    x, y = lum(a), lum(b)
    k = win

    def box(z):
        c = np.cumsum(np.cumsum(np.pad(z, ((1, 0), (1, 0))), 0), 1)
        return (c[k:, k:] - c[:-k, k:] - c[k:, :-k] + c[:-k, :-k]) / (k * k)
    mx, my = box(x), box(y)
    sxx = box(x * x) - mx * mx
    syy = box(y * y) - my * my
    sxy = box(x * y) - mx * my
    c1, c2 = (0.01 * 255) ** 2, (0.03 * 255) ** 2
    s = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sxx + syy + c2))
    return float(s.mean())


def now():
    return lib.ivr_now()


def state_pack(st):
    """STATE as the 184 bytes of the wire (connection.c:167-200)"""
    out = np.zeros(184, np.uint8)
    lib.ivr_state_pack(C.byref(st), _ptr(out))
    return out.tobytes()


def state_unpack(data):
    buf = np.frombuffer(bytes(data), np.uint8).copy()
    st = State()
    lib.ivr_state_unpack(_ptr(buf), C.byref(st))
    return st


def tiff_encode(rgba32, planes=3):
    """writetotiff's bytes (utils.c:18-154) of an (h, w, 4) float32 image"""
    img = np.ascontiguousarray(rgba32, np.float32)
    h, w = img.shape[:2]
    out = np.zeros(512 + w * h * planes, np.uint8)
    n = lib.ivr_tiff_encode(_ptr(img), w, h, planes, _ptr(out), out.size)
    if n < 0:
        raise IvrError("tiff_encode: bad image or planes")
    return out.tobytes()


class Batcher:
    """continuous batching of render requests (ivr_batcher_*)"""

    def __init__(self, ctx, budget_bytes, max_slots, max_width, max_height):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib.ivr_batcher_create(ctx.h, budget_bytes, max_slots, max_width, max_height, C.byref(h)))
        self.h = h
        self._keep = {}

    def add_volume(self, vol):
        rc = lib.ivr_batcher_add_volume(self.h, vol.h)
        if rc < 0:
            check(rc)
        return rc

    def add_surrogate(self, sg, vol=None):
        """serve a trained surrogate (vol: the plain volume of fallback bricks)"""
        rc = lib.ivr_batcher_add_surrogate(self.h, sg.h, vol.h if vol is not None else None)
        if rc < 0:
            check(rc)
        return rc

    def submit(self, rid, volume, desc, lut, out=None):
        """queue a request; returns the host image array it will fill"""
        W, H = desc.state.winwidth, desc.state.winheight
        img = np.empty((H, W, 4), np.uint8) if out is None else out
        lut = np.ascontiguousarray(lut, np.float32)
        r = Request(id=rid, volume=volume, desc=desc, lut=lut.ctypes.data, lutsize=lut.shape[0],
                    rgba8=img.ctypes.data, t_submit=now())
        self._keep[rid] = (img, lut)
        check(lib.ivr_batcher_submit(self.h, C.byref(r)))
        return img

    def busy(self):
        st = self.stats()
        return st["pending"] > 0 or st["in_flight"] > 0

    def progress(self, block=False):
        rc = lib.ivr_batcher_progress(self.h, int(block))
        if rc < 0:
            check(rc)
        return rc

    def poll(self, max_n=256):
        buf = (Completion * max_n)()
        n = lib.ivr_batcher_poll(self.h, buf, max_n)
        out = []
        for k in range(n):
            c = buf[k]
            img, _ = self._keep.pop(c.id, (None, None))
            out.append(dict(id=c.id, status=c.status, t_submit=c.t_submit, t_admit=c.t_admit, t_done=c.t_done,
                            samples=c.samples, image=img))
        return out

    def stats(self):
        st = BatcherStats()
        check(lib.ivr_batcher_stats_get(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in BatcherStats._fields_}

    def close(self):
        if self.h:
            lib.ivr_batcher_destroy(self.h)
            self.h = None
