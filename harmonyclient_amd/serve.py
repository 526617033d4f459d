"""Client of the serve front-end's TCP protocol (the reference's
src/client.cpp and src/master_server.h:5-8), little-endian:

  INSERT phase:  [u8 1][u64 n][n*d f32] -> [u8 status];  [u8 2] -> [u8 status]
  QUERY phase:   [u64 nq][u64 k][nq*d f32] -> [nq*k f32][nq*k i64];  [u64 0][u64 0]
"""
from __future__ import annotations

import socket
import struct

import numpy as np

OP_SHUTDOWN, OP_INSERT, OP_BUILD_DONE = 0, 1, 2


def _recv_exact(sock, n):
    buf = bytearray(n)
    view = memoryview(buf)
    got = 0
    while got < n:
        r = sock.recv_into(view[got:], n - got)
        if r == 0:
            raise ConnectionError("server closed the connection")
        got += r
    return bytes(buf)


class Client:
    def __init__(self, host="127.0.0.1", port=7777, timeout=120.0):
        self.sock = socket.create_connection((host, port), timeout=timeout)
        self.sock.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)

    # -- INSERT phase ------------------------------------------------------
    def insert(self, x) -> int:
        x = np.ascontiguousarray(x, np.float32)
        self.sock.sendall(struct.pack("<BQ", OP_INSERT, x.shape[0]) + x.tobytes())
        return _recv_exact(self.sock, 1)[0]

    def build_done(self) -> int:
        self.sock.sendall(struct.pack("<B", OP_BUILD_DONE))
        return _recv_exact(self.sock, 1)[0]

    # -- QUERY phase -------------------------------------------------------
    def query(self, x, k):
        x = np.ascontiguousarray(x, np.float32)
        nq = x.shape[0]
        self.sock.sendall(struct.pack("<QQ", nq, k) + x.tobytes())
        D = np.frombuffer(_recv_exact(self.sock, nq * k * 4), np.float32).reshape(nq, k)
        I = np.frombuffer(_recv_exact(self.sock, nq * k * 8), np.int64).reshape(nq, k)
        return D, I

    def shutdown_server(self):
        self.sock.sendall(struct.pack("<QQ", 0, 0))

    def close(self):
        self.sock.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
