"""CPU tests of the gpujoind wire protocol as the client speaks it (request layout, frames, mid-stream errors) and of
the daemon's behaviour without a device."""
import os
import socket
import struct
import subprocess
import tempfile
import threading

import pytest

from gsj import client as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DAEMON = os.path.join(ROOT, "postgis-cuda_amd", "bin", "gpujoind")


def test_request_layout_matches_the_reference_protocol():
    b = G.encode_request(G.OP_JOIN, ["a", "b", "1", "all"], token="tok")
    magic, op, nargs = struct.unpack_from("<III", b, 0)
    assert (magic, op, nargs) == (0x52505347, 6, 5)   # "GSPR", JOIN, token + 4 args
    off, args = 12, []
    for _ in range(nargs):
        (n,) = struct.unpack_from("<I", b, off)
        args.append(b[off + 4:off + 4 + n].decode())
        off += 4 + n
    assert args == ["tok", "a", "b", "1", "all"] and off == len(b)


def test_window_is_an_additive_trailing_argument():
    from gsj.client import _window_arg

    assert _window_arg(None) == []
    (w,) = _window_arg((-95.5, 38.5, -95.25, 38.75))
    assert w == "window=-95.5,38.5,-95.25,38.75"
    # floats round-trip exactly through the text form
    x = 0.1 + 0.2
    assert float(_window_arg((x, 0, 1, 1))[0][7:].split(",")[0]) == x


def _serve_once(payload: bytes):
    d = tempfile.mkdtemp()
    path = os.path.join(d, "s.sock")
    srv = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
    srv.bind(path)
    srv.listen(1)

    def run():
        c, _ = srv.accept()
        hdr = c.recv(12)
        _, _, nargs = struct.unpack("<III", hdr)
        for _ in range(nargs):
            (n,) = struct.unpack("<I", c.recv(4))
            while n:
                n -= len(c.recv(n))
        c.sendall(payload)
        c.close()
        srv.close()

    threading.Thread(target=run, daemon=True).start()
    return path


def _row(a, b, v):
    return struct.pack("<H", len(a)) + a.encode() + struct.pack("<H", len(b)) + b.encode() + struct.pack("<d", v)


def test_frames_and_stats_are_parsed():
    stats = b'{"rows": 3}'
    payload = (struct.pack("<I", 0) + struct.pack("<I", 2) + _row("1", "x", 0.5) + _row("2", "y", 1.5) + struct.pack("<I", 1) + _row("3", "z", 2.5)
               + struct.pack("<I", 0) + struct.pack("<I", len(stats)) + stats)
    rows, st, lat = G.Client(_serve_once(payload)).join("a", "b", want_area=True)
    assert rows == [("1", "x", 0.5), ("2", "y", 1.5), ("3", "z", 2.5)] and st == {"rows": 3} and lat >= 0


def test_errors_before_and_during_the_stream():
    msg = b"no such layer: a"
    with pytest.raises(G.DaemonError, match="no such layer"):
        G.Client(_serve_once(struct.pack("<I", 1) + struct.pack("<I", len(msg)) + msg)).join("a", "b")
    mid = b"request: cancelled"
    payload = struct.pack("<I", 0) + struct.pack("<I", 1) + _row("1", "x", 0.0) + struct.pack("<I", 0xFFFFFFFF) + struct.pack("<I", len(mid)) + mid
    with pytest.raises(G.DaemonError, match="cancelled"):
        G.Client(_serve_once(payload)).join("a", "b")


def test_daemon_fails_loudly_without_a_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    if not os.path.exists(DAEMON):
        pytest.skip("daemon not built")
    r = subprocess.run([DAEMON, "--socket", os.path.join(tempfile.mkdtemp(), "d.sock")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "device" in r.stderr
    r = subprocess.run([DAEMON, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "usage" in r.stderr
