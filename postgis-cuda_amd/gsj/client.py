"""Client of the gpujoind wire protocol (the format of the reference's daemon/protocol.h, kept byte for byte):
what the gpu_spatial extension sends and reads, in Python. Used by the daemon tests and the serving benchmark, where
request latency is measured here, from the last request byte to the end-of-stream stats frame."""
from __future__ import annotations

import json
import socket
import struct
import time

GSP_MAGIC = 0x52505347
GSP_OK, GSP_ERR, GSP_FRAME_ERR = 0, 1, 0xFFFFFFFF
OP_LOAD, OP_LOAD_QUERY, OP_DROP, OP_LIST, OP_STATUS, OP_JOIN, OP_CLUSTER, OP_RELATE, OP_DWITHIN = range(1, 10)


class DaemonError(RuntimeError):
    pass


def encode_request(op: int, args, token: str = "") -> bytes:
    parts = [struct.pack("<III", GSP_MAGIC, op, len(args) + 1)]
    for a in [token] + [str(x) for x in args]:
        b = a.encode()
        parts.append(struct.pack("<I", len(b)) + b)
    return b"".join(parts)


class _Reader:
    def __init__(self, sock):
        self.sock = sock
        self.buf = bytearray()

    def take(self, n: int) -> bytes:
        while len(self.buf) < n:
            chunk = self.sock.recv(max(1 << 20, n - len(self.buf)))
            if not chunk:
                raise DaemonError("connection closed")
            self.buf += chunk
        out = bytes(self.buf[:n])
        del self.buf[:n]
        return out

    def u32(self) -> int:
        return struct.unpack("<I", self.take(4))[0]

    def string(self) -> str:
        return self.take(self.u32()).decode()


def _window_arg(window):
    """the optional trailing query-window argument: probe only geometries whose box meets (x0, y0, x1, y1)"""
    return [] if window is None else ["window=" + ",".join(repr(float(v)) for v in window)]


def _row_parser(op: int):
    def ids(r):
        la = struct.unpack("<H", r.take(2))[0]
        a = r.take(la).decode()
        lb = struct.unpack("<H", r.take(2))[0]
        return a, r.take(lb).decode()

    if op in (OP_JOIN, OP_DWITHIN):
        return lambda r: ids(r) + (struct.unpack("<d", r.take(8))[0],)
    if op == OP_RELATE:
        return lambda r: ids(r) + (r.take(9).decode(),)
    if op == OP_CLUSTER:
        def cl(r):
            ln = struct.unpack("<H", r.take(2))[0]
            return r.take(ln).decode(), struct.unpack("<I", r.take(4))[0]
        return cl
    raise ValueError(op)


class Client:
    """One request per connection, as the extension does."""

    def __init__(self, path: str | None = None, tcp: tuple[str, int] | None = None, token: str = "", timeout: float = 600.0):
        self.path, self.tcp, self.token, self.timeout = path, tcp, token, timeout

    def _connect(self):
        if self.tcp:
            s = socket.create_connection(self.tcp, timeout=self.timeout)
        else:
            s = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
            s.settimeout(self.timeout)
            s.connect(self.path)
        return s

    def call(self, op: int, *args):
        """JSON ops: the parsed JSON. Streaming ops: (rows, stats, latency_s)."""
        with self._connect() as s:
            s.sendall(encode_request(op, args, self.token))
            t0 = time.perf_counter()
            r = _Reader(s)
            status = r.u32()
            if status == GSP_ERR:
                raise DaemonError(r.string())
            if status != GSP_OK:
                raise DaemonError(f"bad status {status}")
            if op in (OP_LOAD, OP_LOAD_QUERY, OP_DROP, OP_LIST, OP_STATUS):
                return json.loads(r.string())
            parse = _row_parser(op)
            rows = []
            while True:
                n = r.u32()
                if n == GSP_FRAME_ERR:
                    raise DaemonError(r.string())
                if n == 0:
                    stats = json.loads(r.string())
                    return rows, stats, time.perf_counter() - t0
                rows.extend(parse(r) for _ in range(n))

    # conveniences mirroring the extension's SQL functions
    def load(self, layer, source, subdivide=0, dedup=True):
        return self.call(OP_LOAD, layer, source, "id", "geom", "", subdivide, "true" if dedup else "false")

    def join(self, a, b, want_area=False, self_pairs="all", stats=False, window=None):
        return self.call(OP_JOIN, a, b, "1" if want_area else "0", self_pairs, *(["stats"] if stats else []), *_window_arg(window))

    def relate(self, a, b, pattern="", self_pairs="all", window=None):
        return self.call(OP_RELATE, a, b, pattern, self_pairs, *_window_arg(window))

    def dwithin(self, a, b, mode, dist_a, dist_b, self_pairs="all", extra=None, window=None):
        return self.call(OP_DWITHIN, a, b, mode, dist_a, dist_b, self_pairs, *([extra] if extra else []), *_window_arg(window))

    def cluster(self, layer, window=None):
        return self.call(OP_CLUSTER, layer, *_window_arg(window))
