import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhivf.so)")
    config.addinivalue_line("markers", "slow: full-size property checks")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def w1():
    """Workload 1 (10k x 128 integer mixture, nlist 100) with k-means
    centroids trained by the oracle from the seeded initial centroids."""
    import oracle
    from harmonyclient_amd import datasets as ds

    oracle.build()
    xb, xq = ds.generate(1)
    sample, init = ds.kmeans_init(xb, 100)
    cent = oracle.kmeans(oracle.METRIC_L2, sample, init, 10)
    ix = oracle.Index.build(oracle.METRIC_L2, xb, cent)
    return {"xb": xb, "xq": xq, "cent": cent, "init": init, "sample": sample, "ix": ix}


def brute_topk(x, q, k, metric=1, rows=None, ids=None):
    """Exact top-k by (distance, id) in FP64 (numpy), over `rows` of x."""
    x = np.asarray(x, np.float64)
    q = np.asarray(q, np.float64)
    if rows is not None:
        x = x[rows]
    if ids is None:
        ids = np.arange(x.shape[0]) if rows is None else np.asarray(rows)
    if metric == 1:
        dist = ((x - q) ** 2).sum(1)
        order = np.lexsort((ids, dist))
    else:
        dist = x @ q
        order = np.lexsort((ids, -dist))
    order = order[:k]
    return dist[order], np.asarray(ids)[order]


def random_partials(seed, parts, nq, k_in, metric=1, ties=True):
    """Per-partition sorted top-k_in rows with disjoint ids, some rows short
    (sentinel padded), integer values when `ties` (many equal distances)."""
    rng = np.random.default_rng(seed)
    Dp = np.empty((parts, nq, k_in), np.float32)
    Ip = np.empty((parts, nq, k_in), np.int64)
    fmax = np.finfo(np.float32).max
    sent = fmax if metric == 1 else -fmax
    for p in range(parts):
        for q in range(nq):
            m = int(rng.integers(0, k_in + 1))
            v = (rng.integers(0, 40, m) if ties else rng.random(m) * 100).astype(np.float32)
            ids = (rng.choice(10_000, m, replace=False) * parts + p).astype(np.int64)
            o = np.lexsort((ids, v if metric == 1 else -v))
            Dp[p, q] = sent
            Ip[p, q] = -1
            Dp[p, q, :m] = v[o]
            Ip[p, q, :m] = ids[o]
    return Dp, Ip


@pytest.fixture(scope="session")
def w2s():
    """Workload-2 recipe at 200k rows, nlist 256, 300 queries."""
    import oracle
    from harmonyclient_amd import datasets
    from harmonyclient_amd import index as hx

    print("[case] build 200k x 128 slice (GPU k-means + add)", flush=True)
    w = datasets.WORKLOADS[2]
    xb, xq = datasets.generate(w, nb=200_000, nq=300, wid=2)
    sample, init = datasets.kmeans_init(xb, 256)
    idx = hx.Index(128, 256, 16, hx.METRIC_L2)
    idx.train(sample.shape[0], sample, init_centroids=init, niter=4)
    idx.add(xb.shape[0], xb)
    off, ids, codes = idx.csr()
    ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
    return {"xb": xb, "xq": xq, "idx": idx, "ix": ix}


