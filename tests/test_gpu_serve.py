"""GPU: the serve front-end in front of an index's captured search.  Replies
over the reference's TCP protocol equal the oracle's results, for requests
that are cut into several fixed-size batches, and after an INSERT phase that
built the index through the server (k-means and add on the GPU)."""
import numpy as np
import pytest

import oracle
from harmonyclient_amd import datasets as ds
from harmonyclient_amd import index as hx
from harmonyclient_amd.serve import Client

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _same(D, I, OD, OI):
    assert np.array_equal(I, OI), f"{(I != OI).sum()} id mismatches"
    assert np.array_equal(D.view(np.uint32), OD.view(np.uint32))


def test_serve_queries_match_oracle(w2s):
    _case("serve: 200k slice, batches of 128, learned-free batched search, 300 + 50 queries")
    idx, ix, xq = w2s["idx"], w2s["ix"], w2s["xq"]
    params = hx.make_serve_params(batch=128, search=hx.make_params(nprobe=16, algo=hx.ALGO_BATCHED))
    srv = hx.Server(idx, params=params)
    try:
        with Client(port=srv.port) as c:
            D, I = c.query(xq[:300], 100)
            OD, OI, _ = ix.search(xq[:300], 16, 100)
            _same(D, I, OD, OI)
            D, I = c.query(xq[:50], 10)
            OD, OI, _ = ix.search(xq[:50], 16, 10)
            _same(D, I, OD, OI)
            c.shutdown_server()
        assert srv.wait() == 0, srv.error()
        st = srv.stats()
        print(f"[serve] {st}", flush=True)
        assert st["batches"] == 3 + 1 and st["queries"] == 350
    finally:
        srv.stop()


def test_serve_insert_phase_builds_the_index():
    _case("serve: INSERT phase (4 batches), BUILD_DONE trains and adds on the GPU, then queries")
    xb, xq = ds.generate(1, nb=20000, nq=200)
    idx = hx.Index(128, 64, 8, hx.METRIC_L2)
    params = hx.make_serve_params(batch=64, insert_phase=True, train_niter=4,
                                  search=hx.make_params(nprobe=8, algo=hx.ALGO_BATCHED))
    srv = hx.Server(idx, params=params)
    try:
        with Client(port=srv.port) as c:
            for part in np.array_split(xb, 4):
                assert c.insert(part) == 0
            assert c.build_done() == 0
        assert idx.ntotal == xb.shape[0]
        off, ids, codes = idx.csr()
        ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
        with Client(port=srv.port) as c:
            D, I = c.query(xq, 10)
            OD, OI, _ = ix.search(xq, 8, 10)
            _same(D, I, OD, OI)
            c.shutdown_server()
        assert srv.wait() == 0, srv.error()
    finally:
        srv.stop()
