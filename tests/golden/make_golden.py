"""Write tests/golden/workload1_golden.npz.

Workload 1 (10k x 128 integer mixture, nlist 100, 100 queries, nprobe 10,
k 10): the probe lists and top-10 results of the reference's plain search,
run through oracle/_ref (the reference's own src/heap.hpp driven by the
reference's search loop), plus the warm-up bound (L=10, W=k).  Before
writing, every vector is cross-checked against the oracle and an FP64 brute
force over the probed rows.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from harmonyclient_amd import datasets as ds  # noqa: E402


def main():
    oracle.build()
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref is not built (needs the reference tree)")
    xb, xq = ds.generate(1)
    sample, init = ds.kmeans_init(xb, 100)
    cent = oracle.kmeans(oracle.METRIC_L2, sample, init, 10)
    ix = oracle.Index.build(oracle.METRIC_L2, xb, cent)
    D, I, probes = ix.ref_search(xq, 10, 10)
    OD, OI, _ = ix.search(xq, 10, 10)
    assert np.array_equal(D, OD) and np.array_equal(I, OI)
    for i in range(xq.shape[0]):
        rows = np.concatenate([np.arange(ix.list_off[l], ix.list_off[l + 1]) for l in probes[i]])
        dist = ((ix.codes[rows].astype(np.float64) - xq[i].astype(np.float64)) ** 2).sum(1)
        order = np.lexsort((ix.ids[rows], dist))[:10]
        assert np.array_equal(ix.ids[rows][order], I[i])
        assert np.array_equal(dist[order], D[i].astype(np.float64))
    tau = ix.warmup_tau(xq, probes, 10, 10, 10)
    out = os.path.join(ROOT, "tests", "golden", "workload1_golden.npz")
    np.savez(out, D=D, I=I, probes=probes, tau=tau, centroids=cent,
             base_sha256=np.frombuffer(hashlib.sha256(xb.tobytes()).digest(), np.uint8))
    print("wrote", out)


if __name__ == "__main__":
    main()
