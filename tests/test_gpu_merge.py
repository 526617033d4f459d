"""GPU top-k merge across partitions (src/Index.cpp:1327-1343) against the
oracle's merge, bit for bit, and the list-sharding identity on the GPU:
searching disjoint shards of the lists with the same probes and merging
gives exactly the search over all lists."""
import numpy as np
import pytest

import oracle
from conftest import random_partials
from harmonyclient_amd import index as hx

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _same(D, I, OD, OI):
    assert np.array_equal(I, OI), f"{(I != OI).sum()} id mismatches"
    assert np.array_equal(D.view(np.uint32), OD.view(np.uint32))


@pytest.mark.parametrize("metric", [hx.METRIC_L2, hx.METRIC_IP])
@pytest.mark.parametrize("parts,n,k_in,k", [(1, 50, 10, 10), (4, 100, 100, 100), (8, 64, 128, 1024),
                                            (16, 20, 100, 1), (3, 40, 1024, 100), (5, 30, 7, 33)])
def test_merge_matches_oracle(metric, parts, n, k_in, k):
    _case(f"merge metric={metric} parts={parts} n={n} k_in={k_in} k={k}")
    for ties in (True, False):
        Dp, Ip = random_partials(parts * 1000 + k, parts, n, k_in, metric, ties=ties)
        D, I = hx.merge_topk(metric, Dp, Ip, k)
        OD, OI = oracle.merge(metric, Dp, Ip, k)
        _same(D, I, OD, OI)


def test_merge_edges():
    _case("merge edges: empty batch, all padding, inf/NaN, bad arguments")
    D, I = hx.merge_topk(hx.METRIC_L2, np.zeros((2, 0, 5), np.float32), np.zeros((2, 0, 5), np.int64), 3)
    assert D.shape == (0, 3)
    Dp = np.full((3, 4, 6), FMAX, np.float32)
    Ip = np.full((3, 4, 6), -1, np.int64)
    Dp[2, 1, :3] = [np.inf, np.nan, 4.0]
    Ip[2, 1, :3] = [10, 11, 12]
    D, I = hx.merge_topk(hx.METRIC_L2, Dp, Ip, 5)
    OD, OI = oracle.merge(hx.METRIC_L2, Dp, Ip, 5)
    _same(D, I, OD, OI)
    assert I[1].tolist() == [12, -1, -1, -1, -1] and (I[[0, 2, 3]] == -1).all()
    with pytest.raises(hx.HivfError):
        hx.merge_topk(hx.METRIC_L2, Dp, Ip, 2000)  # k > 1024
    Ip[0, 0, 0] = 1 << 40
    with pytest.raises(hx.HivfError):
        hx.merge_topk(hx.METRIC_L2, Dp, Ip, 5)  # id beyond 32 bits


def test_merge_device_buffers():
    import torch

    _case("merge on device buffers")
    Dp, Ip = random_partials(77, 6, 200, 100)
    dp, ip = torch.from_numpy(Dp).cuda(), torch.from_numpy(Ip).cuda()
    D = torch.empty((200, 100), dtype=torch.float32, device="cuda")
    I = torch.empty((200, 100), dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    hx.merge_topk_device(hx.METRIC_L2, 200, 6, 100, dp.data_ptr(), ip.data_ptr(), 100, D.data_ptr(),
                         I.data_ptr())
    torch.cuda.synchronize()
    OD, OI = oracle.merge(hx.METRIC_L2, Dp, Ip, 100)
    _same(D.cpu().numpy(), I.cpu().numpy(), OD, OI)


@pytest.mark.parametrize("algo", [hx.ALGO_PLAIN, hx.ALGO_BATCHED])
def test_list_sharded_search(w2s, algo):
    """Four shard indexes (list l on shard l % 4, same centroids), each
    searched with the same nprobe: the merged rows equal the full search."""
    _case(f"list-sharded search over 4 shards, algo={algo}")
    idx, xq = w2s["idx"], w2s["xq"]
    off, ids, codes = idx.csr()
    nlist, shards = idx.nlist, 4
    p = hx.make_params(nprobe=16, algo=algo)
    parts_d, parts_i = [], []
    for s in range(shards):
        own = (np.arange(nlist) % shards) == s
        sizes = np.diff(off) * own
        soff = np.concatenate([[0], np.cumsum(sizes)])
        rows = np.concatenate([np.arange(off[l], off[l + 1]) for l in range(nlist) if own[l]])
        sh = hx.Index(idx.d, nlist, 16, hx.METRIC_L2)
        sh.set_centroids(idx.centroids)
        sh.set_lists(soff, ids[rows], codes[rows])
        D, I, _ = sh.search(300, xq, 100, params=p)
        parts_d.append(D)
        parts_i.append(I)
    MD, MI = hx.merge_topk(hx.METRIC_L2, np.stack(parts_d), np.stack(parts_i), 100)
    FD, FI, _ = idx.search(300, xq, 100, params=p)
    _same(MD, MI, FD, FI)
