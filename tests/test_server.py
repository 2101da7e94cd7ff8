"""The steering server (ivr_amd/ivrserver) over the reference frame
protocol: concurrent clients in --listen mode (plain STATE and the versioned
extension) and the drop-in --relay mode behind a relay that behaves like the
reference's myserver (myserver.c:104-165).  Every returned frame equals the
in-process render of the same request."""
import ctypes as C
import os
import socket
import struct
import subprocess
import threading
import time

import numpy as np
import pytest

import ivr_amd as ivr
from ivr_amd import configs

pytestmark = pytest.mark.gpu
SERVER = os.path.join(os.path.dirname(ivr.LIB_PATH), "ivrserver")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def free_port():
    s = socket.socket()
    s.bind(("127.0.0.1", 0))
    p = s.getsockname()[1]
    s.close()
    return p


def recv_all(sock, n):
    buf = bytearray()
    while len(buf) < n:
        chunk = sock.recv(n - len(buf))
        if not chunk:
            raise ConnectionError("peer closed")
        buf += chunk
    return bytes(buf)


def start_server(args, until="ivrserver: ready"):
    """start ivrserver and read its log up to the line containing `until`"""
    p = subprocess.Popen([SERVER] + args, stderr=subprocess.PIPE, text=True)
    t0 = time.time()
    lines = []
    while time.time() - t0 < 120:
        line = p.stderr.readline()
        lines.append(line)
        if until in line:
            threading.Thread(target=lambda: [None for _ in p.stderr], daemon=True).start()
            return p
        if p.poll() is not None:
            break
    p.kill()
    raise RuntimeError("server did not start: " + "".join(lines))


@pytest.fixture(scope="module")
def refs(ctx):
    """the served volumes, generated in-process the same way"""
    a = ctx.volume(64, 64, 64, ivr.IVR_F32, brick=64)
    a.generate(configs.gen_params(1))
    b = ctx.volume(96, 96, 96, ivr.IVR_F32, brick=64)
    b.generate(configs.gen_params(2))
    yield [a, b]
    a.close()
    b.close()


def render_ref(ctx, vol, st, lut):
    tf = ctx.tf(lut)
    fr = ctx.frame(st.winwidth, st.winheight)
    ctx.render(vol, tf, ivr.config_desc(st, skip=True), fr)
    img = fr.download()
    fr.close()
    tf.close()
    return img


def ext_message(rid, volume, lut, st):
    lut = np.ascontiguousarray(lut, np.float32)
    return (b"IVRX" + struct.pack("<IQii", 1, rid, volume, lut.shape[0]) + lut.tobytes() + ivr.state_pack(st))


def test_listen_mode_concurrent_clients(ctx, refs):
    port = free_port()
    srv = start_server(["--listen", str(port), "--volume", "gen:1:64", "--volume", "gen:2:96", "--slots", "6",
                        "--max", "320", "320", "--budget-gb", "2"])
    results, errors = {}, []

    def client(cid):
        try:
            s = socket.create_connection(("127.0.0.1", port))
            assert recv_all(s, 13) == b"Hello, world!"
            s.sendall(b"Hello, server!")
            rng = np.random.default_rng(cid)
            for k in range(4):
                w, h = [(96, 64), (160, 120), (320, 240), (200, 200)][(cid + k) % 4]
                st = ivr.orbit_state(int(rng.integers(1 << 30)), w, h)
                if k % 2 == 0:  # the plain reference protocol: volume 0, default LUT
                    s.sendall(ivr.state_pack(st))
                    img = np.frombuffer(recv_all(s, w * h * 4), np.uint8).reshape(h, w, 4)
                    results[(cid, k)] = (0, st, configs.tf_lut(1), img)
                else:  # extension: volume 1, jittered LUT, request id echoed
                    lut = ivr.tf_build(2002, 0.05, 0.05, int(rng.integers(1 << 30)))
                    rid = cid * 100 + k
                    s.sendall(ext_message(rid, 1, lut, st))
                    hdr = recv_all(s, 24)
                    assert hdr[:4] == b"IVRR"
                    got_id, status, gw, gh = struct.unpack("<Qiii", hdr[4:])
                    assert (got_id, status, gw, gh) == (rid, 0, w, h)
                    img = np.frombuffer(recv_all(s, w * h * 4), np.uint8).reshape(h, w, 4)
                    results[(cid, k)] = (1, st, lut, img)
            fin = ivr.State()
            fin.finish = 1
            s.sendall(ivr.state_pack(fin))
            s.close()
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=client, args=(c,)) for c in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    srv.terminate()
    srv.wait(timeout=30)
    assert not errors, errors
    assert len(results) == 12
    for key, (vol, st, lut, img) in results.items():
        assert np.array_equal(img, render_ref(ctx, refs[vol], st, lut)), key


def test_raw_big_endian_volume(ctx, tmp_path):
    """--volume raw: a big-endian float32 file of arbitrary dims and value
    range (as readvoldata reads the reference's datasets, main.c:629-706),
    LUT spanning the data range; frames equal the in-process renders"""
    nx, ny, nz = 61, 47, 53
    blobs = ivr.generate_host(configs.gen_params(1), nx, ny, nz)
    host = (np.float32(1468.0) + np.float32(3518.0) * blobs).astype(np.float32)
    path = str(tmp_path / "shell.raw")
    host.astype(">f4").tofile(path)
    port = free_port()
    srv = start_server(["--listen", str(port), "--volume", "raw:%s:%d:%d:%d:0" % (path, nx, ny, nz),
                        "--max", "200", "200", "--budget-gb", "1"])
    try:
        lo, hi = np.float32(host.min()), np.float32(host.max())
        scale = np.float32(255.0) / (hi - lo)
        vol = ctx.volume(nx, ny, nz, ivr.IVR_F32, brick=64)
        vol.upload(host)
        tf = ctx.tf(configs.tf_lut(2))
        s = socket.create_connection(("127.0.0.1", port))
        assert recv_all(s, 13) == b"Hello, world!"
        s.sendall(b"Hello, server!")
        for seed in (3201, 3202):
            st = ivr.orbit_state(seed, 150, 110)
            s.sendall(ivr.state_pack(st))
            img = np.frombuffer(recv_all(s, 150 * 110 * 4), np.uint8).reshape(110, 150, 4)
            d = ivr.config_desc(st, skip=True, lutscale=float(scale))
            d.lutoffset = int(np.float32(lo * scale))
            fr = ctx.frame(150, 110)
            ctx.render(vol, tf, d, fr)
            ref = fr.download()
            fr.close()
            assert np.array_equal(img, ref), seed
            assert (ref[..., 3] > 0).mean() > 0.05
        fin = ivr.State()
        fin.finish = 1
        s.sendall(ivr.state_pack(fin))
        s.close()
        tf.close()
        vol.close()
    finally:
        srv.terminate()
        srv.wait(timeout=30)


RELAY = os.path.join(ROOT, "oracle", "_ref", "myserver")


@pytest.mark.skipif(not os.path.exists(RELAY), reason="reference relay not built (make -C oracle relay)")
def test_behind_the_reference_relay(ctx, refs):
    """the reference's own relay (myserver.c, compiled from its sources)
    between a client and ivrserver: client first, render server second, as
    the reference deploys them; frames equal the in-process renders"""
    port = free_port()
    relay = subprocess.Popen([RELAY, str(port)], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    srv = None
    try:
        cl = None
        t0 = time.time()
        while time.time() - t0 < 30:
            try:
                cl = socket.create_connection(("127.0.0.1", port), timeout=60)
                break
            except OSError:
                time.sleep(0.1)
        assert cl is not None, "relay did not listen"
        assert recv_all(cl, 13) == b"Hello, world!"
        cl.sendall(b"Hello, server!")
        srv = start_server(["--relay", "127.0.0.1:%d" % port, "--volume", "gen:1:64", "--max", "256", "256",
                            "--budget-gb", "1"], until="ivrserver: ready")
        for seed in (3101, 3102, 3103):
            st = ivr.orbit_state(seed, 112, 80)
            cl.sendall(ivr.state_pack(st))
            img = np.frombuffer(recv_all(cl, 112 * 80 * 4), np.uint8).reshape(80, 112, 4)
            assert np.array_equal(img, render_ref(ctx, refs[0], st, configs.tf_lut(1))), seed
        fin = ivr.State()
        fin.finish = 1
        cl.sendall(ivr.state_pack(fin))
        assert srv.wait(timeout=60) == 0
        cl.close()
    finally:
        relay.kill()  # the reference relay serves until it is stopped
        relay.wait(timeout=30)
        if srv is not None and srv.poll() is None:
            srv.kill()


def test_relay_mode_is_a_dropin_render_server(ctx, refs):
    """fake myserver: accept the GPU server, then a client; relay STATE and frames"""
    port = free_port()
    ls = socket.socket()
    ls.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
    ls.bind(("127.0.0.1", port))
    ls.listen(2)
    # the server connects to the relay once its volumes are resident and
    # waits for the relay's greeting before it reports ready
    srv = start_server(["--relay", "127.0.0.1:%d" % port, "--volume", "gen:1:64", "--max", "256", "256",
                        "--budget-gb", "1"], until="volume 0 ready")
    ls.settimeout(60)
    conn, _ = ls.accept()
    conn.settimeout(60)
    conn.sendall(b"Hello, world!")
    assert recv_all(conn, 14) == b"Hello, server!"
    for seed in (3001, 3002):
        st = ivr.orbit_state(seed, 128, 96)
        conn.sendall(ivr.state_pack(st))
        img = np.frombuffer(recv_all(conn, 128 * 96 * 4), np.uint8).reshape(96, 128, 4)
        assert np.array_equal(img, render_ref(ctx, refs[0], st, configs.tf_lut(1)))
    fin = ivr.State()
    fin.finish = 1
    conn.sendall(ivr.state_pack(fin))
    assert srv.wait(timeout=60) == 0
    conn.close()
    ls.close()


def test_requests_sent_ahead_are_served(ctx, refs):
    """a client that sends its next requests before the reply to the first
    (two STATEs and two extension messages in one sendall, then a finish in
    the same segment as the last request) gets every frame in order, and the
    finish buffered behind a request ends the relay session once it is served"""
    port = free_port()
    srv = start_server(["--listen", str(port), "--volume", "gen:1:64", "--volume", "gen:2:96", "--max", "256",
                        "256", "--budget-gb", "1"])
    try:
        s = socket.create_connection(("127.0.0.1", port), timeout=60)
        assert recv_all(s, 13) == b"Hello, world!"
        s.sendall(b"Hello, server!")
        sts = [ivr.orbit_state(seed, w, h) for seed, (w, h) in zip((3301, 3302), ((120, 88), (96, 72)))]
        s.sendall(ivr.state_pack(sts[0]) + ivr.state_pack(sts[1]))
        for st in sts:
            w, h = st.winwidth, st.winheight
            img = np.frombuffer(recv_all(s, w * h * 4), np.uint8).reshape(h, w, 4)
            assert np.array_equal(img, render_ref(ctx, refs[0], st, configs.tf_lut(1)))
        lut = ivr.tf_build(2002, 0.05, 0.05, 77)
        s.sendall(ext_message(11, 1, lut, sts[1]) + ext_message(12, 1, lut, sts[0]))
        for rid, st in ((11, sts[1]), (12, sts[0])):
            hdr = recv_all(s, 24)
            assert hdr[:4] == b"IVRR" and struct.unpack("<Qiii", hdr[4:]) == (rid, 0, st.winwidth, st.winheight)
            img = np.frombuffer(recv_all(s, st.winwidth * st.winheight * 4), np.uint8)
            assert np.array_equal(img.reshape(st.winheight, st.winwidth, 4), render_ref(ctx, refs[1], st, lut))
        s.close()
    finally:
        srv.terminate()
        srv.wait(timeout=30)
    # relay role: the request and the finish arrive in one segment
    port = free_port()
    ls = socket.socket()
    ls.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
    ls.bind(("127.0.0.1", port))
    ls.listen(1)
    srv = start_server(["--relay", "127.0.0.1:%d" % port, "--volume", "gen:1:64", "--max", "256", "256",
                        "--budget-gb", "1"], until="volume 0 ready")
    try:
        ls.settimeout(60)
        conn, _ = ls.accept()
        conn.settimeout(60)
        conn.sendall(b"Hello, world!")
        assert recv_all(conn, 14) == b"Hello, server!"
        fin = ivr.State()
        fin.finish = 1
        conn.sendall(ivr.state_pack(sts[0]) + ivr.state_pack(fin))
        w, h = sts[0].winwidth, sts[0].winheight
        img = np.frombuffer(recv_all(conn, w * h * 4), np.uint8).reshape(h, w, 4)
        assert np.array_equal(img, render_ref(ctx, refs[0], sts[0], configs.tf_lut(1)))
        assert srv.wait(timeout=60) == 0
        conn.close()
    finally:
        ls.close()
        if srv.poll() is None:
            srv.kill()


@pytest.mark.skipif(not os.path.exists(RELAY), reason="reference relay not built (make -C oracle relay)")
@pytest.mark.parametrize("cmap,xfer", [(4, 2), (2, 3)])
def test_reference_command_line_behind_the_reference_relay(ctx, tmp_path, cmap, xfer):
    """ivrserver started with the reference server's own command line (raw
    volume, colormap, transfer function, decomposition and view arguments)
    behind the reference relay compiled from myserver.c: every frame is
    bit-identical to the compiled reference's createlocalimage for the same
    STATE and the LUT its initvolluts builds (maxlutindx = hsize - 1 as the
    MPI render server sets it), quantised as main.c:1404 does"""
    import oracle
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("the compiled reference (oracle/_ref) is not built")
    nx, ny, nz = 61, 47, 53
    blobs = ivr.generate_host(configs.gen_params(1), nx, ny, nz)
    host = (np.float32(1468.0) + np.float32(3518.0) * blobs).astype(np.float32)
    path = str(tmp_path / "shell.raw")
    host.tofile(path)  # little endian: ENDIAN 1
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    tile = 8
    assert ref.ivrref_init(path.encode(), nx, ny, nz, cmap, xfer, tile) == 0
    hs = int(np.float32(np.float32(host.max() - host.min()) + np.float32(1.0)))
    lut = np.zeros((hs, 4), np.float32)
    ref.ivrref_get_lut(f(lut), hs)
    ref.ivrref_set_lut(f(lut), hs, ref.ivrref_lutoffset(), hs - 1)
    ref.ivrref_set_xfer(xfer)
    port = free_port()
    relay = subprocess.Popen([RELAY, str(port)], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    srv = None
    log_path = str(tmp_path / "server.log")
    try:
        cl = None
        t0 = time.time()
        while time.time() - t0 < 30:
            try:
                cl = socket.create_connection(("127.0.0.1", port), timeout=30)
                break
            except OSError:
                time.sleep(0.1)
        assert cl is not None, "relay did not listen"
        assert recv_all(cl, 13) == b"Hello, world!"
        cl.sendall(b"Hello, server!")
        env = dict(os.environ, IVR_RELAY="127.0.0.1:%d" % port)
        args = [path, str(nx), str(ny), str(nz), "1", str(cmap), str(xfer), "1", "1", "1", str(tile), str(tile),
                "180", "1", "0", "0"]
        with open(log_path, "w") as log:
            srv = subprocess.Popen([SERVER] + args, stderr=log, env=env)
        # the last window is not a whole number of tiles: both render the viewport setviewport
        # rounds it to (196x100 -> 192x96); the reference would send only its bytes, fewer than
        # the relay forwards, and stall; ivrserver sends the window with the viewport at its
        # bottom left and zeros beyond
        for seed, (w, h) in ((3401, (160, 120)), (3402, (128, 128)), (3403, (196, 100))):
            st = ivr.orbit_state(seed, w, h)
            vw, vh = w // tile * tile, h // tile * tile
            cl.sendall(ivr.state_pack(st))
            win = np.frombuffer(recv_all(cl, w * h * 4), np.uint8).reshape(h, w, 4)
            img = win[:vh, :vw]
            assert not win[vh:].any() and not win[:, vw:].any()
            out = np.zeros((h, w, 4), np.float32)
            rect = (C.c_int * 4)()
            ref.ivrref_render(f(np.array(st.mvm, np.float32)), f(np.array(st.pm, np.float32)),
                              f(np.array(st.vv, np.float32)), f(np.array(st.newconst, np.float32)),
                              C.c_float(st.alpha), w, h, f(out), rect)
            want = oracle.quantize(out.reshape(-1, 4)[:vw * vh].reshape(vh, vw, 4))
            assert (want[..., 3] > 0).mean() > 0.05
            assert np.array_equal(img, want), (seed, int(np.abs(img.astype(int) - want.astype(int)).max()))
        fin = ivr.State()
        fin.finish = 1
        cl.sendall(ivr.state_pack(fin))
        assert srv.wait(timeout=60) == 0
        cl.close()
    except Exception:
        print("server log:\n" + open(log_path).read(), flush=True)
        raise
    finally:
        relay.kill()
        relay.wait(timeout=30)
        if srv is not None and srv.poll() is None:
            srv.kill()
