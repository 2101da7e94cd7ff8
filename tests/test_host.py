"""CPU-only checks of libivr.so: it loads, exports every entry point the
public headers declare, and its host-side code (generator, transfer-function
builder, steering states) behaves as specified.  No GPU call is made."""
import ctypes as C
import glob
import hashlib
import json
import math
import os
import re

import numpy as np
import pytest

import ivr_amd as ivr
from ivr_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        if h.endswith("ivr_compat.h"):  # libivr_compat.so, checked in test_compat.py
            continue
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 25, names
    lib = C.CDLL(ivr.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_host_generator_deterministic_and_pinned():
    gp = configs.gen_params(1)
    gp.value_scale = 255.0
    a = ivr.generate_host(gp, 64, 64, 64)
    b = ivr.generate_host(gp, 64, 64, 64)
    assert np.array_equal(a, b)
    meta = json.load(open(os.path.join(ROOT, "tests", "golden", "golden_meta.json")))
    assert hashlib.sha256(a.tobytes()).hexdigest() == meta["sha256"]
    assert a.min() == 0.0 and a.max() == 255.0


def test_generator_exp_and_noise_properties():
    # blobs only, one blob at the centre: the field is a radial Gaussian
    gp = ivr.gen_params(ivr.GEN_BLOBS, seed=7, nblobs=1)
    v = ivr.generate_host(gp, 33, 33, 33)
    c = np.unravel_index(np.argmax(v), v.shape)
    assert v[c] == 1.0 and v.min() == 0.0
    # noise with a quantile threshold zeroes that fraction
    gp = ivr.gen_params(ivr.GEN_NOISE_BLOBS, seed=9, nblobs=4, empty_frac=0.7)
    v = ivr.generate_host(gp, 40, 40, 40)
    assert abs(np.mean(v == 0.0) - 0.7) < 0.002 and v.max() == 1.0


def test_tf_builder_closed_form():
    lut = ivr.tf_build(2002, 0.05)
    assert lut.shape == (256, 4)
    assert lut[0, 3] == 0.0
    assert (lut >= 0).all() and (lut[:, :3] <= 1).all() and (lut[:, 3] <= 0.05 + 1e-7).all()
    # piecewise linear with at most 3 interior knots: second differences vanish
    # except near the knots
    d2 = np.abs(np.diff(lut.astype(np.float64), 2, axis=0)).max(axis=1)
    assert (d2 > 1e-5).sum() <= 6
    j = ivr.tf_build(2002, 0.05, jitter=0.05, jitter_seed=5)
    assert np.abs(j - lut).max() <= 0.05 + 0.05 * 255 / 1.0  # positions move, values stay bounded
    assert j[0, 3] == 0.0


def test_orbit_state_is_the_clients_modelview():
    st = ivr.orbit_state(3002, 1024, 1024)
    m = np.array(st.mvm, np.float64).reshape(4, 4).T  # column-major -> rows
    R = m[:3, :3]
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
    assert np.allclose(m[:3, 3], [0, 0, -5]) and np.allclose(m[3], [0, 0, 0, 1])
    t = math.tan(math.radians(43.6028) / 2) * 5
    assert np.allclose(st.vv, [-t, t, -t, t], rtol=1e-6) and abs(t - 2.0) < 1e-4
    p = np.array(st.pm, np.float64).reshape(4, 4).T
    assert abs(p[1, 1] - 1 / math.tan(math.radians(43.6028) / 2)) < 1e-5 and p[3, 2] == -1.0


def test_state_layout():
    assert C.sizeof(ivr.State) == 184
    assert ivr.State.mvm.offset == 56 and ivr.State.pm.offset == 120


def test_state_float_swap_is_recvstate_under_bigendian():
    """a STATE whose floats arrive in the other byte order (a BIGENDIAN
    build's peer, connection.c:191-197) unpacks to the sender's STATE after
    ivr_state_swap_floats; swapping twice is the identity"""
    st = ivr.orbit_state(77, 320, 200)
    st.alpha = 0.625
    st.newconst[2] = -0.25
    wire = bytearray(ivr.state_pack(st))
    for off in range(12, 184, 4):  # every float of the wire message, byte-reversed
        wire[off:off + 4] = wire[off:off + 4][::-1]
    got = ivr.state_unpack(bytes(wire))
    assert bytes(got) != bytes(st)
    ivr.lib.ivr_state_swap_floats(C.byref(got))
    assert bytes(got) == bytes(st)
    ivr.lib.ivr_state_swap_floats(C.byref(got))
    ivr.lib.ivr_state_swap_floats(C.byref(got))
    assert bytes(got) == bytes(st)


@pytest.mark.parametrize("planes", [1, 3])
def test_tiff_bytes_match_the_reference_writetotiff(planes):
    """ivr_tiff_encode writes the bytes of the reference's writetotiff
    (utils.c:18-154, compiled from its sources) for the same RGBA32F image"""
    import oracle
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("the compiled reference (oracle/_ref) is not built")
    rng = np.random.default_rng(planes)
    h, w = 37, 53
    img = rng.uniform(0.0, 1.0, (h, w, 4)).astype(np.float32)
    img[0, 0] = [0.0, 1.0, 0.5, 1.0]
    want = np.zeros(512 + w * h * planes, np.uint8)
    ref.writetotiff(want.ctypes.data_as(C.c_char_p), C.c_long(w), C.c_long(h),
                    C.c_long(planes), img.ctypes.data_as(C.POINTER(C.c_float)))
    got = ivr.tiff_encode(img, planes)
    assert got[:4] == b"MM\x00*" and len(got) == want.size
    assert got == want.tobytes()
