"""GPU: hivf_index_search_device_noalloc (the search entry point that
allocates no device memory) through the C ABI against
hivf_index_search_device on the same inputs: identical ids and distance
bits, on two calls in a row, the second with a smaller input.  Both the
plain and the pruned scan (warm-up bound), an index whose d is not a
multiple of 4 (the padded-query staging in the scratch array), and the
rejection of sizes beyond the array."""
import ctypes as C

import numpy as np
import pytest

from harmonyclient_amd import datasets as ds
from harmonyclient_amd import index as hx

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _noalloc(idx, n, q, k, D, I, params):
    return hx.lib().hivf_index_search_device_noalloc(idx._h, n, q.data_ptr(), k, D.data_ptr(), I.data_ptr(),
                                                    C.byref(params), None)


def _check(idx, xq, k, params, sizes):
    import torch

    for n in sizes:
        q = torch.from_numpy(np.ascontiguousarray(xq[:n])).cuda()
        D0 = torch.empty((n, k), dtype=torch.float32, device="cuda")
        I0 = torch.empty((n, k), dtype=torch.int64, device="cuda")
        D1 = torch.full((n, k), -7.0, dtype=torch.float32, device="cuda")
        I1 = torch.full((n, k), -7, dtype=torch.int64, device="cuda")
        idx.search_device(n, q.data_ptr(), k, D0.data_ptr(), I0.data_ptr(), params=params, stats=False)
        assert _noalloc(idx, n, q, k, D1, I1, params) == 0, hx.lib().hivf_last_error(idx._h)
        torch.cuda.synchronize()
        assert torch.equal(I0, I1), f"n={n}: ids differ"
        assert torch.equal(D0.view(torch.int32), D1.view(torch.int32)), f"n={n}: distance bits differ"
        print(f"[noalloc] n={n} k={k} algo={params.algo}: identical", flush=True)


@pytest.mark.parametrize("algo,tau", [(hx.ALGO_PLAIN, hx.TAU_NONE), (hx.ALGO_PRUNED, hx.TAU_WARMUP)])
def test_noalloc_matches_search_device(w2s, algo, tau):
    _case(f"noalloc search algo={algo} on the 200k slice, 300 then 120 queries")
    _check(w2s["idx"], w2s["xq"], 100, hx.make_params(nprobe=16, algo=algo, tau=tau), [300, 120])


def test_noalloc_padded_dimension():
    _case("noalloc search, d = 30 (queries staged in the scratch array), 500 then 77 queries")
    rng = np.random.default_rng(4)
    xb = rng.integers(0, 40, (20000, 30)).astype(np.float32)
    xq = rng.integers(0, 40, (500, 30)).astype(np.float32)
    idx = hx.Index(30, 50, 8, hx.METRIC_L2)
    idx.train(xb.shape[0], xb, niter=3)
    idx.add(xb.shape[0], xb)
    for algo in (hx.ALGO_PLAIN, hx.ALGO_PRUNED):
        _check(idx, xq, 10, hx.make_params(nprobe=8, algo=algo), [500, 77])


def test_noalloc_rejects_sizes_beyond_its_scratch(w2s):
    import torch

    _case("noalloc search: n * nprobe beyond the scratch array is an error, not a fault")
    idx = w2s["idx"]
    n = 70000  # above 65536 queries
    q = torch.zeros((n, 128), dtype=torch.float32, device="cuda")
    D = torch.empty((n, 1), dtype=torch.float32, device="cuda")
    I = torch.empty((n, 1), dtype=torch.int64, device="cuda")
    rc = _noalloc(idx, n, q, 1, D, I, hx.make_params(nprobe=16))
    assert rc == -1 and b"at most" in hx.lib().hivf_last_error(idx._h)
    n = 5000  # 5000 x 256 probes = 1.28M pairs, above 2^20
    rc = _noalloc(idx, n, q[:n], 1, D[:n], I[:n], hx.make_params(nprobe=256))
    assert rc == -1
    assert _noalloc(idx, 10, q[:10], 1, D[:10], I[:10], hx.make_params(nprobe=16, algo=hx.ALGO_BATCHED)) == -5
