"""The serve front-end's transport and batching on the CPU: a server whose
batch search is a Python brute force (hivf_server_start_fn) must answer the
reference's protocol exactly -- INSERT batches, BUILD_DONE, query requests
of any size cut into fixed batches, several k, a malformed request closing
only its own connection, and the shutdown header.  The same server code
drives the index's captured search on the GPU (test_gpu_serve.py)."""
import socket
import struct
import threading

import numpy as np
import pytest

from conftest import brute_topk
from harmonyclient_amd import index as hx
from harmonyclient_amd.serve import Client


class BruteForce:
    def __init__(self, d):
        self.d = d
        self.x = np.zeros((0, d), np.float32)
        self.calls = []
        self.lock = threading.Lock()

    def build(self, x):
        self.x = np.array(x, np.float32)

    def search(self, q, k):
        with self.lock:
            self.calls.append(q.shape[0])
        D = np.full((q.shape[0], k), np.finfo(np.float32).max, np.float32)
        I = np.full((q.shape[0], k), -1, np.int64)
        for i in range(q.shape[0]):
            dist, ids = brute_topk(self.x, q[i], k)
            D[i, :len(ids)] = dist
            I[i, :len(ids)] = ids
        return D, I


def closed_by_server(s):
    """The server closed this connection (an orderly close, or a reset when
    unread request bytes were pending on its side)."""
    try:
        return s.recv(16) == b""
    except ConnectionResetError:
        return True


def expected(bf, q, k):
    D = np.full((q.shape[0], k), np.finfo(np.float32).max, np.float32)
    I = np.full((q.shape[0], k), -1, np.int64)
    for i in range(q.shape[0]):
        dist, ids = brute_topk(bf.x, q[i], k)
        D[i, :len(ids)] = dist
        I[i, :len(ids)] = ids
    return D, I


def test_insert_then_query_in_fixed_batches():
    d, batch = 8, 32
    rng = np.random.default_rng(0)
    base = rng.integers(0, 50, (300, d)).astype(np.float32)
    bf = BruteForce(d)
    srv = hx.Server(d=d, search_fn=bf.search, build_fn=bf.build, batch=batch, insert_phase=True)
    try:
        with Client(port=srv.port) as c:
            assert c.insert(base[:100]) == 0
            assert c.insert(base[100:]) == 0
            assert c.build_done() == 0
        assert np.array_equal(bf.x, base)
        q = rng.integers(0, 50, (75, d)).astype(np.float32)
        with Client(port=srv.port) as c:
            for nq, k in [(75, 10), (1, 3), (32, 400), (64, 1)]:
                D, I = c.query(q[:nq] if nq <= 75 else q, k)
                ED, EI = expected(bf, q[:nq], k)
                assert np.array_equal(I, EI) and np.array_equal(D.view(np.uint32), ED.view(np.uint32))
            c.shutdown_server()
        assert srv.wait() == 0
        st = srv.stats()
        # batches of at most 32 rows: 75 -> 32 + 32 + 11, then 1, 32, 32 + 32
        assert bf.calls == [32, 32, 11, 1, 32, 32, 32]
        assert st["requests"] == 4 and st["queries"] == 75 + 1 + 32 + 64 and st["batches"] == 7
        assert st["rows_inserted"] == 300 and st["connections"] == 2
    finally:
        srv.stop()


def test_malformed_request_closes_only_its_connection():
    d = 4
    bf = BruteForce(d)
    bf.x = np.arange(40, dtype=np.float32).reshape(10, d)
    srv = hx.Server(d=d, search_fn=bf.search, batch=8)
    try:
        s = socket.create_connection(("127.0.0.1", srv.port), timeout=30)
        s.sendall(struct.pack("<QQ", 2, 0) + b"\0" * 32)  # k = 0
        assert closed_by_server(s)
        s.close()
        assert "k outside" in srv.error()
        s = socket.create_connection(("127.0.0.1", srv.port), timeout=30)
        s.sendall(struct.pack("<QQ", 1, 5000))  # k above 1024
        assert closed_by_server(s)
        s.close()
        with Client(port=srv.port) as c:  # the server still answers
            q = np.ones((3, d), np.float32)
            D, I = c.query(q, 2)
            ED, EI = expected(bf, q, 2)
            assert np.array_equal(I, EI)
            c.shutdown_server()
        srv.wait()
    finally:
        srv.stop()


def test_unknown_opcode_and_failed_search_report_errors():
    d = 4
    bf = BruteForce(d)
    srv = hx.Server(d=d, search_fn=lambda q, k: (_ for _ in ()).throw(RuntimeError("no")), build_fn=bf.build,
                    batch=8, insert_phase=True)
    try:
        with Client(port=srv.port) as c:
            c.sock.sendall(struct.pack("<B", 9))
            assert c.sock.recv(1) == b"\x01"  # unknown opcode: status error
            assert c.insert(np.ones((4, d), np.float32)) == 0
            assert c.build_done() == 0
        with Client(port=srv.port) as c:
            c.sock.sendall(struct.pack("<QQ", 1, 2) + np.ones(d, np.float32).tobytes())
            assert closed_by_server(c.sock)  # the search failed: connection closed
        assert "search failed" in srv.error()
    finally:
        srv.stop()


def test_stop_unblocks_a_waiting_server():
    srv = hx.Server(d=4, search_fn=lambda q, k: None, batch=8)
    port = srv.port
    assert port > 0
    srv.stop()  # no client ever connected: the listener is closed, the thread joins
    with pytest.raises(OSError):
        socket.create_connection(("127.0.0.1", port), timeout=5)
