"""GPU checks of the captured search (hipGraph replay) and the tile-shape
autotuner: both must give exactly the eager search's results.

A captured search is replayed several times with other eager searches on the
same handle (which overwrite its scratch) and host synchronisation between
the replays; every replay is compared with the eager result of the queries
the buffer holds at that moment.
"""
import numpy as np
import pytest

from harmonyclient_amd import index as hx

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _same(D, I, OD, OI):
    assert np.array_equal(I, OI), f"{(I != OI).sum()} id mismatches"
    assert np.array_equal(D.view(np.uint32), OD.view(np.uint32))


@pytest.mark.parametrize("algo,tau", [(hx.ALGO_PLAIN, hx.TAU_NONE), (hx.ALGO_PRUNED, hx.TAU_WARMUP),
                                      (hx.ALGO_BATCHED, hx.TAU_WARMUP)])
def test_graph_replay_matches_eager(w2s, algo, tau):
    import torch

    _case(f"graph replay algo={algo} tau={tau} on the 200k slice, 300 queries, k=100")
    idx, xq = w2s["idx"], w2s["xq"]
    n, k = 300, 100
    rng = np.random.default_rng(5)
    q = torch.from_numpy(xq).cuda()
    D = torch.empty((n, k), dtype=torch.float32, device="cuda")
    I = torch.empty((n, k), dtype=torch.int64, device="cuda")
    pg = hx.make_params(nprobe=16, algo=algo, tau=tau, use_graph=True)
    pe = hx.make_params(nprobe=16, algo=algo, tau=tau)
    batches = [xq] + [np.ascontiguousarray(xq[rng.permutation(n)]) for _ in range(3)]
    for r, xb in enumerate(batches):
        q.copy_(torch.from_numpy(xb))
        torch.cuda.synchronize()
        idx.search_device(n, q.data_ptr(), k, D.data_ptr(), I.data_ptr(), params=pg, stats=False)
        torch.cuda.synchronize()
        got_d, got_i = D.cpu().numpy(), I.cpu().numpy()
        # eager work on the same handle between replays (other shape: k=10)
        idx.search(50, xq[::-1][:50], 10, params=pe)
        OD, OI, _ = idx.search(n, xb, k, params=pe)
        _same(got_d, got_i, OD, OI)
        print(f"[graph] replay {r}: identical to eager", flush=True)


def test_graph_recaptures_on_new_shape(w2s):
    import torch

    _case("graph mode: a new k or buffer recaptures")
    idx, xq = w2s["idx"], w2s["xq"]
    q = torch.from_numpy(xq).cuda()
    pe = hx.make_params(nprobe=16, algo=hx.ALGO_BATCHED)
    pg = hx.make_params(nprobe=16, algo=hx.ALGO_BATCHED, use_graph=True)
    for k in (100, 10, 100):
        D = torch.empty((300, k), dtype=torch.float32, device="cuda")
        I = torch.empty((300, k), dtype=torch.int64, device="cuda")
        for _ in range(2):
            idx.search_device(300, q.data_ptr(), k, D.data_ptr(), I.data_ptr(), params=pg, stats=False)
            torch.cuda.synchronize()
            OD, OI, _ = idx.search(300, xq, k, params=pe)
            _same(D.cpu().numpy(), I.cpu().numpy(), OD, OI)


def test_autotune_keeps_results(w2s):
    import torch

    _case("autotune the batched tile shape on the 200k slice")
    idx, xq = w2s["idx"], w2s["xq"]
    pe = hx.make_params(nprobe=16, algo=hx.ALGO_BATCHED)
    OD, OI, _ = idx.search(300, xq, 100, params=pe)
    q = torch.from_numpy(xq).cuda()
    torch.cuda.synchronize()
    r = idx.autotune(300, q.data_ptr(), 100, params=pe, reps=2)
    print(f"[tune] {r}", flush=True)
    # integer data, d = 128: the int8 scan takes 32 or 64 queries per item
    assert r["tried"] == 16 and r["batch_queries"] in (32, 64)
    assert r["batch_rows"] in (256, 512, 1024, 2048)
    assert r["batch_waves"] in (8, 16) and 0 < r["ms"] <= r["ms_default"] * 1.5
    D, I, st = idx.search(300, xq, 100, params=pe)
    assert (st.batch_queries, st.batch_rows, st.batch_waves) == (r["batch_queries"], r["batch_rows"],
                                                                   r["batch_waves"])
    _same(D, I, OD, OI)
    idx.set_tune()  # back to the built-in default for the other tests
    _, _, st = idx.search(300, xq, 100, params=pe)
    # (integer data, d = 128: the int8 scan's default is 64-query items)
    assert (st.batch_queries, st.batch_rows, st.batch_waves) == (64, 1024, 16)


def test_graph_replay_after_buffer_growth(w2s):
    """A captured search must not replay on buffers a later, larger call on
    the same handle freed and reallocated (scratch, staging, coarse and
    batched buffers, the autotuner's shape change)."""
    import torch

    _case("graph replay after a larger eager search and an autotune on the same handle")
    idx, xq, xb = w2s["idx"], w2s["xq"], w2s["xb"]
    n, k = 300, 100
    pg = hx.make_params(nprobe=16, algo=hx.ALGO_BATCHED, use_graph=True)
    pe = hx.make_params(nprobe=16, algo=hx.ALGO_BATCHED)
    q = torch.from_numpy(xq).cuda()
    D = torch.empty((n, k), dtype=torch.float32, device="cuda")
    I = torch.empty((n, k), dtype=torch.int64, device="cuda")
    for _ in range(2):  # capture, then one replay
        idx.search_device(n, q.data_ptr(), k, D.data_ptr(), I.data_ptr(), params=pg, stats=False)
    torch.cuda.synchronize()
    OD, OI, _ = idx.search(n, xq, k, params=pe)
    _same(D.cpu().numpy(), I.cpu().numpy(), OD, OI)
    # larger eager searches on the same handle grow (free and reallocate)
    # every scratch buffer the captured search used
    big = np.ascontiguousarray(xb[:4000] + 1.0)
    for algo in (hx.ALGO_BATCHED, hx.ALGO_PLAIN, hx.ALGO_PRUNED):
        idx.search(4000, big, 200, params=hx.make_params(nprobe=64, algo=algo))
    qb = torch.from_numpy(big).cuda()
    torch.cuda.synchronize()
    r = idx.autotune(4000, qb.data_ptr(), 200, params=hx.make_params(nprobe=64, algo=hx.ALGO_BATCHED), reps=1)
    print(f"[tune] {r}", flush=True)
    D.fill_(0.0)
    I.fill_(0)
    for rep in range(3):
        idx.search_device(n, q.data_ptr(), k, D.data_ptr(), I.data_ptr(), params=pg, stats=False)
        torch.cuda.synchronize()
        _same(D.cpu().numpy(), I.cpu().numpy(), OD, OI)
        print(f"[graph] replay {rep} after growth: identical to eager", flush=True)
    idx.set_tune()
