"""GPU: the batched search with per-query probe counts read from device
memory (nprobe_per_query_dev) and the learned bound, eager and replayed as a
captured graph over batches whose queries and probe counts change between
replays, against the plain kernel given the same counts in host memory, and
a sample against the oracle.  The data: Zipf-popular integer mixture (skewed
list sizes), probe counts uniform in [8, 256]."""
import numpy as np
import pytest

from harmonyclient_amd import datasets as ds
from harmonyclient_amd import index as hx

pytestmark = pytest.mark.gpu

NB, NLIST, K, B, NBATCH = 300_000, 1024, 100, 1024, 3


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _diff(tag, D, I, PD, PI, npq):
    bad = np.where((I != PI).any(1) | (D.view(np.uint32) != PD.view(np.uint32)).any(1))[0]
    if bad.size:
        q = bad[0]
        j = np.where((I[q] != PI[q]) | (D[q].view(np.uint32) != PD[q].view(np.uint32)))[0][0]
        print(f"[{tag}] {bad.size} queries differ, first {q} (nprobe {npq[q]}) at rank {j}: "
              f"got ({D[q, j]!r}, {I[q, j]}) want ({PD[q, j]!r}, {PI[q, j]})", flush=True)
    return bad.size


@pytest.fixture(scope="module")
def zipf(oracle_lib):
    import oracle

    _case(f"build {NB} x 128 Zipf integer mixture, nlist {NLIST}")
    w = ds.WORKLOADS[5]
    xb, xq = ds.generate(w, nb=NB, nq=B * NBATCH, wid=5)
    idx = hx.Index(128, NLIST, 64, hx.METRIC_L2)
    idx.train(xb.shape[0], xb, niter=4)
    idx.add(xb.shape[0], xb)
    _case("bound model on 3000 traces of separate queries (nprobe 132)")
    xt = ds.generate(w, nb=0, nq=3000, wid=5, query_seed=3005)[1]
    _, st = idx.train_tau_model(xt, K, 132, epochs=20, lr=2e-3, quantile=0.9, miss_rate=0.01)
    print("[train]", st, flush=True)
    npq = np.random.default_rng(11).integers(8, 257, B * NBATCH).astype(np.int32)
    off, ids, codes = idx.csr()
    ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
    return {"idx": idx, "xq": xq, "npq": npq, "ix": ix}


def _plain(idx, xq, npq):
    D, I, _ = idx.search(xq.shape[0], xq, K, params=hx.make_params(nprobe=256, algo=hx.ALGO_PLAIN,
                                                                    nprobe_per_query=npq))
    return D, I


def test_batched_host_counts_match_plain(zipf):
    _case("batched, learned bound, per-query counts from host memory")
    idx, xq, npq = zipf["idx"], zipf["xq"][:B], zipf["npq"][:B]
    PD, PI = _plain(idx, xq, npq)
    D, I, st = idx.search(B, xq, K, params=hx.make_params(nprobe=256, algo=hx.ALGO_BATCHED, tau=hx.TAU_LEARNED,
                                                          nprobe_per_query=npq))
    print(f"[host counts] fallback {st.fallback_queries} (overflow {st.overflow_queries}) "
          f"rescans {st.exact_rescans} survivors/query {st.survivors / B:.0f}", flush=True)
    assert _diff("host counts", D, I, PD, PI, npq) == 0


@pytest.mark.parametrize("graph", [False, True])
def test_batched_device_counts_match_plain(zipf, graph):
    import torch

    _case(f"batched, learned bound, per-query counts from device memory, graph={graph}")
    idx, xq_all, npq_all = zipf["idx"], zipf["xq"], zipf["npq"]
    dev = torch.device("cuda:0")
    qbuf = torch.empty((B, 128), dtype=torch.float32, device=dev)
    nbuf = torch.empty(B, dtype=torch.int32, device=dev)
    Dbuf = torch.empty((B, K), dtype=torch.float32, device=dev)
    Ibuf = torch.empty((B, K), dtype=torch.int64, device=dev)
    prm = hx.make_params(nprobe=256, algo=hx.ALGO_BATCHED, tau=hx.TAU_LEARNED, use_graph=graph,
                         nprobe_per_query_dev=nbuf.data_ptr())
    bad = 0
    for rep in range(2):
        for b in range(NBATCH):
            xq, npq = xq_all[b * B:(b + 1) * B], npq_all[b * B:(b + 1) * B]
            qbuf.copy_(torch.from_numpy(xq))
            nbuf.copy_(torch.from_numpy(npq))
            torch.cuda.synchronize()
            idx.search_device(B, qbuf.data_ptr(), K, Dbuf.data_ptr(), Ibuf.data_ptr(), params=prm, stats=False)
            torch.cuda.synchronize()
            D, I = Dbuf.cpu().numpy(), Ibuf.cpu().numpy()
            PD, PI = _plain(idx, xq, npq)
            nb = _diff(f"device counts graph={graph} pass {rep} batch {b}", D, I, PD, PI, npq)
            bad += nb
            if rep == 0 and b == 0:
                for i in range(40):  # the oracle, one query at a time with its own probe count
                    OD, OI, _ = zipf["ix"].search(xq[i:i + 1], int(npq[i]), K)
                    assert np.array_equal(PI[i:i + 1], OI) and np.array_equal(PD[i:i + 1].view(np.uint32),
                                                                               OD.view(np.uint32))
    assert bad == 0


def test_streamed_batches_on_one_stream(zipf):
    """The serving pattern: every batch's copies in, captured search and
    copies out enqueued back to back on one stream, one synchronisation at
    the end."""
    import torch

    _case("captured batches streamed on one torch stream, device counts, one sync at the end")
    idx, xq_all, npq_all = zipf["idx"], zipf["xq"], zipf["npq"]
    dev = torch.device("cuda:0")
    st = torch.cuda.Stream()
    qall = torch.from_numpy(xq_all).to(dev)
    nall = torch.from_numpy(npq_all).to(dev)
    qbuf = torch.empty((B, 128), dtype=torch.float32, device=dev)
    nbuf = torch.empty(B, dtype=torch.int32, device=dev)
    Dbuf = torch.empty((B, K), dtype=torch.float32, device=dev)
    Ibuf = torch.empty((B, K), dtype=torch.int64, device=dev)
    Dall = torch.empty((B * NBATCH, K), dtype=torch.float32, device=dev)
    Iall = torch.empty((B * NBATCH, K), dtype=torch.int64, device=dev)
    prm = hx.make_params(nprobe=256, algo=hx.ALGO_BATCHED, tau=hx.TAU_LEARNED, use_graph=True,
                         nprobe_per_query_dev=nbuf.data_ptr())
    torch.cuda.synchronize()
    for _ in range(2):
        with torch.cuda.stream(st):
            for b in range(NBATCH):
                qbuf.copy_(qall[b * B:(b + 1) * B])
                nbuf.copy_(nall[b * B:(b + 1) * B])
                idx.search_device(B, qbuf.data_ptr(), K, Dbuf.data_ptr(), Ibuf.data_ptr(), params=prm,
                                  stats=False, stream=st.cuda_stream)
                Dall[b * B:(b + 1) * B].copy_(Dbuf)
                Iall[b * B:(b + 1) * B].copy_(Ibuf)
        torch.cuda.synchronize()
    PD, PI = _plain(idx, xq_all, npq_all)
    assert _diff("streamed", Dall.cpu().numpy(), Iall.cpu().numpy(), PD, PI, npq_all) == 0
