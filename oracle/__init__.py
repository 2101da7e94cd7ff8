"""ctypes binding of the CPU oracle (liboracle.so) and, where it was built,
the reference renderer compiled from its own sources (_ref/libivrref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libivrref.so")

COMPAT, CONFIG = 0, 1


class View(C.Structure):
    _fields_ = [("mvm", C.c_float * 16), ("pm", C.c_float * 16), ("vv", C.c_float * 4),
                ("newconst", C.c_float * 6), ("alpha", C.c_float), ("width", C.c_int),
                ("height", C.c_int)]


class Params(C.Structure):
    _fields_ = [("mode", C.c_int), ("xfer", C.c_int), ("step_vox", C.c_float), ("ert", C.c_float),
                ("lutscale", C.c_float), ("lutoffset", C.c_int), ("maxlutindx", C.c_int),
                ("brick", C.c_int), ("brick_id", C.c_int * 3), ("sort_last", C.c_int)]


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _load():
    if not os.path.exists(LIB):
        raise ImportError("oracle not built: make -C oracle")
    lib = C.CDLL(LIB)
    lib.ivro_render.restype = C.c_longlong
    lib.ivro_render_f64.restype = C.c_longlong
    lib.ivro_sample_ref.restype = C.c_float
    lib.ivro_sample_tri.restype = C.c_float
    lib.ivro_sample_tri64.restype = C.c_double
    lib.ivro_sample_ref.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_float, C.c_float, C.c_float]
    lib.ivro_sample_tri.argtypes = lib.ivro_sample_ref.argtypes
    lib.ivro_sample_nearest.restype = C.c_float
    lib.ivro_sample_nearest.argtypes = lib.ivro_sample_ref.argtypes
    return lib


lib = _load()


def view_of(state):
    """oracle View from an ivr_amd.State (same fields as STATE)."""
    v = View()
    v.mvm[:] = list(state.mvm)
    v.pm[:] = list(state.pm)
    v.vv[:] = list(state.vv)
    v.newconst[:] = list(state.newconst)
    v.alpha = state.alpha
    v.width = state.winwidth
    v.height = state.winheight
    return v


def params(mode=CONFIG, xfer=3, step_vox=0.5, ert=0.99, lutscale=255.0, lutoffset=0,
           maxlutindx=255, brick=0, brick_id=(0, 0, 0), sort_last=False):
    p = Params(mode=mode, xfer=xfer, step_vox=step_vox, ert=ert, lutscale=lutscale,
               lutoffset=lutoffset, maxlutindx=maxlutindx, brick=brick, sort_last=int(sort_last))
    p.brick_id[:] = list(brick_id)
    return p


def render(vol, view, lut, p, rows=None, nthreads=0, f64=False):
    """Render into a (H, W, 4) array (row 0 = bottom); returns (img, samples)."""
    vol = np.ascontiguousarray(vol, np.float32)
    lut = np.ascontiguousarray(lut, np.float32)
    nz, ny, nx = vol.shape
    H, W = view.height, view.width
    r0, r1 = (0, H) if rows is None else rows
    if f64:
        out = np.zeros((H, W, 4), np.float64)
        n = lib.ivro_render_f64(vol.ctypes.data_as(C.c_void_p), nx, ny, nz, C.byref(view), _f(lut),
                                lut.shape[0], C.byref(p), out.ctypes.data_as(C.c_void_p), r0, r1,
                                nthreads)
    else:
        out = np.zeros((H, W, 4), np.float32)
        n = lib.ivro_render(vol.ctypes.data_as(C.c_void_p), nx, ny, nz, C.byref(view), _f(lut),
                            lut.shape[0], C.byref(p), _f(out), r0, r1, nthreads)
    return out, n


def quantize(img, clamp=True):
    img = np.ascontiguousarray(img, np.float32)
    out = np.empty(img.shape, np.uint8)
    lib.ivro_quantize(_f(img), C.c_longlong(img.size), out.ctypes.data_as(C.c_void_p), int(clamp))
    return out


def composite_tiles(tiles):
    """reference compositetiles over tiles[k] (k = 0 nearest), into zeros."""
    tiles = np.ascontiguousarray(tiles, np.float32)
    num = tiles.shape[0]
    npix = tiles[0].size // 4
    dst = np.zeros(tiles.shape[1:], np.float32)
    lib.ivro_composite_tiles(_f(dst), _f(tiles), num, npix)
    return dst


def box(nx, ny, nz):
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    lib.ivro_box(nx, ny, nz, _f(bmin), _f(bmax))
    return bmin, bmax


def firstvoxel(box6, ray):
    """getfirstvoxel (ray.c:333-384): (rc, [x, y, z]); box6 = aabb min then max"""
    b = np.ascontiguousarray(box6, np.float32)
    r = np.ascontiguousarray(ray, np.float32)
    v = (C.c_int * 3)()
    rc = lib.ivro_firstvoxel(_f(b[:3]), _f(b[3:]), _f(r), v)
    return rc, list(v)


def voxelsonray(box6, dims, ray, maxdepth, fill=-7.0):
    """getvoxelsonray (ray.c:387-483): (rc, maxdepth offsets as float32; fill where not written)"""
    b = np.ascontiguousarray(box6, np.float32)
    r = np.ascontiguousarray(ray, np.float32)
    d = (C.c_int * 3)(*[int(x) for x in dims])
    out = np.full(maxdepth, fill, np.float32)
    rc = lib.ivro_voxelsonray(_f(b[:3]), _f(b[3:]), d, _f(r), maxdepth, _f(out))
    return rc, out


def gradient(vol):
    """gradient field (nz, ny, nx, 4) of a (nz, ny, nx) float32 volume"""
    v = np.ascontiguousarray(vol, np.float32)
    out = np.empty(v.shape + (4,), np.float32)
    lib.ivro_gradient(_f(v), v.shape[2], v.shape[1], v.shape[0], _f(out))
    return out


def ref_lib():
    """the reference compiled from its own sources, or None where absent"""
    if not os.path.exists(REF_LIB):
        return None
    return C.CDLL(REF_LIB)


# ---------------------------------------------------------------- surrogate oracle

lib.ivros_level_res.restype = C.c_int
lib.ivros_level_res.argtypes = [C.c_int, C.c_int, C.c_double]
lib.ivros_index.restype = C.c_uint32
lib.ivros_index.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint32, C.c_int, C.c_int, C.c_int]
lib.ivros_encode.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_double, C.c_void_p, C.c_void_p]
lib.ivros_encode_backward.argtypes = [C.c_uint32, C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
lib.ivros_mlp.restype = C.c_double
lib.ivros_mlp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                          C.c_void_p]
lib.ivros_eval.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_int,
                           C.c_int, C.c_void_p]
lib.ivros_adam.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_float, C.c_float,
                           C.c_float, C.c_float, C.c_int]

SG_NPARAM = 64 * 32 + 64 + 64 * 64 + 64 + 64 + 1


def _v(a):
    return C.c_void_p(a.ctypes.data)


def sg_encode(table, T, x, nmin=16, growth=1.5):
    """table float32 [16][T][2]; x float64 [3] -> enc float64 [32]"""
    table = np.ascontiguousarray(table, np.float32)
    x = np.ascontiguousarray(x, np.float64)
    enc = np.zeros(32, np.float64)
    lib.ivros_encode(_v(table), T, nmin, growth, _v(x), _v(enc))
    return enc


def sg_forward_backward(params, table, T, xyz, targets, bf16=True, nmin=16, growth=1.5):
    """MLP + encoding forward/backward over points of one brick; returns
    (y, loss, gparams, gtable) in double precision"""
    xyz = np.ascontiguousarray(xyz, np.float64)
    n = xyz.shape[0]
    enc = np.stack([sg_encode(table, T, p, nmin, growth) for p in xyz])
    params = np.ascontiguousarray(params, np.float32)
    t = np.ascontiguousarray(targets, np.float32)
    y = np.zeros(n, np.float64)
    g = np.zeros(SG_NPARAM, np.float64)
    denc = np.zeros((n, 32), np.float64)
    loss = lib.ivros_mlp(_v(params), _v(enc), _v(t), n, int(bf16), _v(y), _v(g), _v(denc))
    gt = np.zeros((16, T, 2), np.float64)
    for k in range(n):
        lib.ivros_encode_backward(T, nmin, growth, _v(np.ascontiguousarray(xyz[k])),
                                  _v(np.ascontiguousarray(denc[k])), _v(gt))
    return y, loss, g, gt


def sg_eval(params, table, T, xyz, bf16=True, nmin=16, growth=1.5):
    """forward only (encoding + MLP) over points of one brick, OpenMP over
    the points; float64 [n]"""
    xyz = np.ascontiguousarray(xyz, np.float64)
    table = np.ascontiguousarray(table, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    y = np.zeros(xyz.shape[0], np.float64)
    lib.ivros_eval(_v(table), T, nmin, growth, _v(params), _v(xyz), xyz.shape[0], int(bf16), _v(y))
    return y


def sg_adam(p, g, m, v, lr, b1, b2, eps, t):
    p, g, m, v = (np.ascontiguousarray(a, np.float32).copy() for a in (p, g, m, v))
    lib.ivros_adam(_v(p), _v(g), _v(m), _v(v), p.size, lr, b1, b2, eps, t)
    return p, m, v
