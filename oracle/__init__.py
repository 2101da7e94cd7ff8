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
                ("brick", C.c_int), ("brick_id", C.c_int * 3)]


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
           maxlutindx=255, brick=0, brick_id=(0, 0, 0)):
    p = Params(mode=mode, xfer=xfer, step_vox=step_vox, ert=ert, lutscale=lutscale,
               lutoffset=lutoffset, maxlutindx=maxlutindx, brick=brick)
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


def ref_lib():
    """the reference compiled from its own sources, or None where absent"""
    if not os.path.exists(REF_LIB):
        return None
    return C.CDLL(REF_LIB)
