"""GPU parity: libhivf.so (through the C ABI) against the CPU oracle.

Integer, index and id results must be bit-identical; distances too, since
both sides evaluate the same sequential FMA chain.  Each test prints the case
it runs so a failing GPU run names it.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, brute_topk
from harmonyclient_amd import datasets as ds
from harmonyclient_amd import fvecs
from harmonyclient_amd import index as hx

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


def assert_same(D, I, OD, OI):
    assert np.array_equal(I, OI), f"{(I != OI).sum()} id mismatches"
    assert same_bits(D, OD)


@pytest.fixture(scope="module")
def g1(w1):
    _case("build workload-1 index on the GPU")
    idx = hx.Index(128, 100, 10, hx.METRIC_L2)
    idx.set_centroids(w1["cent"])
    idx.add(w1["xb"].shape[0], w1["xb"])
    return idx


def test_coarse_matches_oracle(g1, w1):
    _case("coarse nq=100 nlist=100 P=10")
    d, l = g1.coarse(w1["xq"], 10)
    od, ol = oracle.coarse(oracle.METRIC_L2, w1["xq"], w1["cent"], 10)
    assert np.array_equal(l, ol)
    assert same_bits(d, od)


def test_coarse_all_lists(g1, w1):
    _case("coarse P=nlist (sorted full list)")
    d, l = g1.coarse(w1["xq"], 0)
    od, ol = oracle.coarse(oracle.METRIC_L2, w1["xq"], w1["cent"], 0)
    assert l.shape == (100, 100)
    assert np.array_equal(l, ol) and same_bits(d, od)


@pytest.mark.parametrize("metric", [hx.METRIC_L2, hx.METRIC_IP])
@pytest.mark.parametrize("d,nlist,P", [(128, 256, 16), (128, 1024, 64), (420, 300, 100), (500, 64, 64),
                                       (20, 40, 7), (3, 33, 5)])
def test_coarse_mfma_cases(metric, d, nlist, P):
    """MFMA coarse quantizer (BF16 pieces + exact rescoring near the P-th
    bound) against the oracle's FP32 chain: integer and real-valued data,
    both metrics, duplicated centroids (exact ties resolved by id)."""
    _case(f"coarse MFMA metric={metric} d={d} nlist={nlist} P={P}")
    rng = np.random.default_rng(d * 7 + nlist)
    for kind in ("int", "real"):
        if kind == "int":
            cent = rng.integers(0, 256, (nlist, d)).astype(np.float32)
            q = rng.integers(0, 256, (97, d)).astype(np.float32)
        else:
            cent = (rng.standard_t(3, (nlist, d)) * 10 + 3).astype(np.float32)
            q = (rng.standard_t(3, (97, d)) * 10 + 3).astype(np.float32)
        cent[nlist // 2] = cent[1]      # exact tie between two centroids
        q[5] = cent[3]                  # a query equal to a centroid
        idx = hx.Index(d, nlist, P, metric)
        idx.set_centroids(cent)
        dist, lists = idx.coarse(q, P)
        od, ol = oracle.coarse(metric, q, cent, P)
        assert np.array_equal(lists, ol), f"{kind}: {(lists != ol).sum()} probe mismatches"
        assert same_bits(dist, od)


@pytest.mark.parametrize("metric", [hx.METRIC_L2, hx.METRIC_IP])
def test_coarse_mfma_workgroup_pick(metric):
    """The one-wave pick hands a query to the workgroup pick when nprobe
    exceeds its buffer (600 > 512) or when its candidates do: vectors with a
    large common offset make the error bound wider than the spread of the
    distances, so every centroid is a candidate."""
    _case(f"coarse MFMA workgroup pick metric={metric}")
    rng = np.random.default_rng(11)
    cent = rng.integers(0, 256, (1024, 64)).astype(np.float32)
    q = rng.integers(0, 256, (40, 64)).astype(np.float32)
    idx = hx.Index(64, 1024, 600, metric)
    idx.set_centroids(cent)
    dist, lists = idx.coarse(q, 600)
    od, ol = oracle.coarse(metric, q, cent, 600)
    assert np.array_equal(lists, ol) and same_bits(dist, od)
    off_c = (10000.0 + rng.standard_normal((1024, 64))).astype(np.float32)
    off_q = (10000.0 + rng.standard_normal((40, 64))).astype(np.float32)
    idx.set_centroids(off_c)
    dist, lists = idx.coarse(off_q, 20)
    od, ol = oracle.coarse(metric, off_q, off_c, 20)
    assert np.array_equal(lists, ol) and same_bits(dist, od)


def test_coarse_mfma_overflow():
    """Values so large that the MFMA form's norms overflow: those bounds are
    not finite, the centroids are rescored exactly, and the result still
    equals the oracle (a query equal to a huge centroid finds it at 0)."""
    _case("coarse MFMA with overflowing norms")
    rng = np.random.default_rng(3)
    cent = rng.integers(0, 256, (40, 16)).astype(np.float32)
    q = rng.integers(0, 256, (9, 16)).astype(np.float32)
    cent[7] = 1e20
    q[3] = 1e20
    idx = hx.Index(16, 40, 5, hx.METRIC_L2)
    idx.set_centroids(cent)
    dist, lists = idx.coarse(q, 5)
    od, ol = oracle.coarse(oracle.METRIC_L2, q, cent, 5)
    assert np.array_equal(lists, ol) and same_bits(dist, od)
    assert lists[3, 0] == 7 and dist[3, 0] == 0.0


def test_add_matches_oracle(g1, w1):
    _case("add: list contents and order")
    off, ids, codes = g1.csr()
    ix = w1["ix"]
    assert np.array_equal(off, ix.list_off)
    assert np.array_equal(ids, ix.ids)
    assert np.array_equal(codes, ix.codes)
    assert g1.ntotal == w1["xb"].shape[0]


def test_train_matches_oracle(w1):
    _case("train: 10 Lloyd rounds, GPU assignment vs oracle")
    idx = hx.Index(128, 100, 10, hx.METRIC_L2)
    idx.train(w1["sample"].shape[0], w1["sample"], init_centroids=w1["init"], niter=10)
    assert same_bits(idx.centroids, w1["cent"])


def test_plain_search_matches_oracle(g1, w1):
    _case("plain search workload 1")
    D, I, st = g1.search(100, w1["xq"], 10, params=hx.make_params(nprobe=10, algo=hx.ALGO_PLAIN))
    OD, OI, ost = w1["ix"].search(w1["xq"], 10, 10)
    assert_same(D, I, OD, OI)
    assert st.candidates == ost.candidates


def test_golden_workload1(g1, w1):
    _case("plain search vs committed golden vectors")
    g = np.load(os.path.join(GOLDEN, "workload1_golden.npz"))
    D, I, _ = g1.search(100, w1["xq"], 10, params=hx.make_params(nprobe=10))
    assert np.array_equal(I, g["I"]) and same_bits(D, g["D"])
    _, l = g1.coarse(w1["xq"], 10)
    assert np.array_equal(l, g["probes"])


@pytest.mark.parametrize("tau_src", [hx.TAU_WARMUP, hx.TAU_NONE])
@pytest.mark.parametrize("dim_block", [32, 8, 128, 7, 70, 0])
def test_pruned_matches_plain(g1, w1, tau_src, dim_block):
    _case(f"pruned search tau={tau_src} dim_block={dim_block}")
    p = hx.make_params(nprobe=10, algo=hx.ALGO_PRUNED, tau=tau_src, dim_block=dim_block)
    D, I, st = g1.search(100, w1["xq"], 10, params=p)
    OD, OI, _ = w1["ix"].search(w1["xq"], 10, 10)
    assert_same(D, I, OD, OI)
    if dim_block < 128:
        assert st.blocks_skipped > 0


def test_warmup_tau_matches_oracle(g1, w1):
    _case("warm-up bound L=10 W=k")
    _, probes = oracle.coarse(oracle.METRIC_L2, w1["xq"], w1["cent"], 10)
    for L, W in [(10, 10), (3, 50), (1, 1), (20, 1000)]:
        t = g1.warmup_tau(w1["xq"], probes, 10, L=L, W=W)
        ot = w1["ix"].warmup_tau(w1["xq"], probes, 10, L, W)
        assert same_bits(t, ot), (L, W)


def test_scan_with_oracle_tau(g1, w1):
    _case("scan entry with external tau")
    _, probes = oracle.coarse(oracle.METRIC_L2, w1["xq"], w1["cent"], 10)
    tau = w1["ix"].warmup_tau(w1["xq"], probes, 10, 10, 10)
    D, I, st = g1.scan(w1["xq"], probes, 10, algo=hx.ALGO_PRUNED, tau=tau, dim_block=32)
    OD, OI, ost = w1["ix"].search_pruned(w1["xq"], probes, 10, 32, tau, True)
    assert_same(D, I, OD, OI)


def test_ip_metric(w1):
    _case("IP metric workload 1")
    cent = oracle.kmeans(oracle.METRIC_IP, w1["sample"], w1["init"], 5)
    ix = oracle.Index.build(oracle.METRIC_IP, w1["xb"], cent)
    idx = hx.Index(128, 100, 10, hx.METRIC_IP)
    idx.set_centroids(cent)
    idx.add(w1["xb"].shape[0], w1["xb"])
    off, ids, _ = idx.csr()
    assert np.array_equal(off, ix.list_off) and np.array_equal(ids, ix.ids)
    D, I, _ = idx.search(100, w1["xq"], 10, params=hx.make_params(nprobe=10))
    OD, OI, _ = ix.search(w1["xq"], 10, 10)
    assert_same(D, I, OD, OI)
    assert np.all(np.diff(D, axis=1) <= 0)


def test_ip_train_matches_oracle(w1):
    _case("IP train")
    idx = hx.Index(128, 100, 10, hx.METRIC_IP)
    idx.train(w1["sample"].shape[0], w1["sample"], init_centroids=w1["init"], niter=5)
    cent = oracle.kmeans(oracle.METRIC_IP, w1["sample"], w1["init"], 5)
    assert same_bits(idx.centroids, cent)


def test_edge_k_exceeds_candidates(g1, w1):
    _case("k larger than the probed rows: sentinel tail")
    D, I, _ = g1.search(5, w1["xq"][:5], 1024, params=hx.make_params(nprobe=1))
    OD, OI, _ = w1["ix"].search(w1["xq"][:5], 1, 1024)
    assert_same(D, I, OD, OI)
    assert (I == -1).any() and np.all(D[I == -1] == np.finfo(np.float32).max)


def test_edge_empty_batch_and_bad_args(g1, w1):
    _case("n = 0 and argument errors")
    D, I, st = g1.search(0, w1["xq"][:0], 10)
    assert D.shape == (0, 10)
    with pytest.raises(hx.HivfError):
        g1.search(1, w1["xq"][:1], 0)
    with pytest.raises(hx.HivfError):
        g1.search(1, w1["xq"][:1], 2048)


def test_edge_empty_lists_and_tiny_dims():
    _case("reference fixtures: d=20 (5 rows) and iris d=4")
    root = os.path.join(GOLDEN, "reference_fixtures")
    for base, query, lead, nlist in [("test_base.fvecs", "test_query.fvecs", True, 3),
                                     ("iris_base.fvecs", "iris_query.fvecs", False, 7)]:
        xb = fvecs.read_fvecs(os.path.join(root, base), leading_word=lead)
        xq = fvecs.read_fvecs(os.path.join(root, query), leading_word=lead)
        cent = xb[:nlist].copy()
        cent[-1] += 1000.0  # a far centroid: its list stays empty
        ix = oracle.Index.build(oracle.METRIC_L2, xb, cent)
        idx = hx.Index(xb.shape[1], nlist, nlist, hx.METRIC_L2)
        idx.set_centroids(cent)
        idx.add(xb.shape[0], xb)
        assert idx.list_size(nlist - 1) == 0
        for k in (1, 3, 200):
            D, I, _ = idx.search(xq.shape[0], xq, k, params=hx.make_params(nprobe=nlist))
            OD, OI, _ = ix.search(xq, nlist, k)
            assert_same(D, I, OD, OI)
            P, Pi, _ = idx.search(xq.shape[0], xq, k, params=hx.make_params(
                nprobe=nlist, algo=hx.ALGO_PRUNED, dim_block=4))
            assert_same(P, Pi, OD, OI)


def test_sift10k_reference_queries(g1, w1):
    _case("sift10k queries (reference fixture) against workload-1 base")
    xq = fvecs.read_fvecs(os.path.join(GOLDEN, "reference_fixtures", "sift10k_query.fvecs"))
    D, I, _ = g1.search(xq.shape[0], xq, 10, params=hx.make_params(nprobe=10))
    OD, OI, _ = w1["ix"].search(xq, 10, 10)
    assert_same(D, I, OD, OI)


def test_save_load_roundtrip(g1, w1, tmp_path):
    _case("save_index / load_index")
    path = str(tmp_path / "w1.index")
    g1.save_index(path)
    idx = hx.Index.load_index(path, nprobe=10)
    assert (idx.d, idx.nlist, idx.metric) == (128, 100, hx.METRIC_L2)
    off, ids, codes = idx.csr()
    assert np.array_equal(off, w1["ix"].list_off) and np.array_equal(ids, w1["ix"].ids)
    D, I, _ = idx.search(100, w1["xq"], 10)
    OD, OI, _ = w1["ix"].search(w1["xq"], 10, 10)
    assert_same(D, I, OD, OI)


def test_load_reference_layout_file(w1, tmp_path):
    """A file written in the reference's layout by an independent writer
    (tests/reffile.py: src/Index.cpp:1708-1732, src/IVF.cpp:76-99) -- lists
    out of order, one empty list left out, OPT_TRIANGLE, OPT_SUBNN_L2 and
    OPT_SUBNN_IP extras that must be skipped, and a list id repeated (the
    later record replaces the earlier, as load_index does) -- loads through
    the C ABI and searches like the oracle on the same lists.  The library's
    own save writes the same bytes back for an extras-free file."""
    import reffile

    _case("load a reference-layout index file with extras, an empty list and a repeated list")
    ix = w1["ix"]
    d, nlist = 128, 100
    lists = {l: (ix.ids[ix.list_off[l]:ix.list_off[l + 1]], ix.codes[ix.list_off[l]:ix.list_off[l + 1]])
             for l in range(nlist)}
    empty = int(np.argmax([len(lists[l][0]) for l in range(nlist)]))
    lists[empty] = (lists[empty][0][:0], lists[empty][1][:0])  # list emptied: not written
    path = tmp_path / "ref.index"
    with open(path, "wb") as f:
        f.write(reffile.header(d, nlist, 1, w1["cent"]))
        for j, l in enumerate(reversed(range(nlist))):
            ids, codes = lists[l]
            if len(ids) == 0:
                continue
            opt = [reffile.OPT_NONE, reffile.OPT_TRIANGLE, reffile.OPT_SUBNN_L2,
                   reffile.OPT_SUBNN_IP | reffile.OPT_TRIANGLE][j % 4]
            if l == 7:  # a stale record of list 7 first, replaced by the real one
                f.write(reffile.list_record(l, ids[:3], codes[:3] + 1.0, opt=opt, sub_k=2))
            f.write(reffile.list_record(l, ids, codes, opt=opt, sub_k=2, seed=l))
    idx = hx.Index.load_index(str(path), nprobe=10)
    off, ids, codes = idx.csr()
    want_off = np.zeros(nlist + 1, np.int64)
    want_off[1:] = np.cumsum([len(lists[l][0]) for l in range(nlist)])
    assert np.array_equal(off, want_off)
    assert np.array_equal(ids, np.concatenate([lists[l][0] for l in range(nlist)]))
    assert np.array_equal(codes, np.concatenate([lists[l][1] for l in range(nlist)]))
    oix = oracle.Index(oracle.METRIC_L2, w1["cent"], off, ids, codes)
    for algo in (hx.ALGO_PLAIN, hx.ALGO_BATCHED):
        D, I, _ = idx.search(100, w1["xq"], 10, params=hx.make_params(nprobe=10, algo=algo))
        OD, OI, _ = oix.search(w1["xq"], 10, 10)
        assert_same(D, I, OD, OI)
    # an extras-free file in the reference's layout comes back byte for byte
    plain = tmp_path / "plain.index"
    reffile.write_index(plain, d, nlist, 1, w1["cent"], lists, sub_k=0)
    again = tmp_path / "again.index"
    hx.Index.load_index(str(plain)).save_index(str(again))
    assert plain.read_bytes() == again.read_bytes()


def test_search_device_buffers(g1, w1):
    _case("search on device buffers (torch plumbing)")
    import torch

    q = torch.from_numpy(w1["xq"]).cuda()
    D = torch.empty((100, 10), dtype=torch.float32, device="cuda")
    I = torch.empty((100, 10), dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    g1.search_device(100, q.data_ptr(), 10, D.data_ptr(), I.data_ptr(),
                     params=hx.make_params(nprobe=10))
    torch.cuda.synchronize()
    OD, OI, _ = w1["ix"].search(w1["xq"], 10, 10)
    assert_same(D.cpu().numpy(), I.cpu().numpy(), OD, OI)


@pytest.mark.parametrize("algo", [hx.ALGO_PLAIN, hx.ALGO_PRUNED])
def test_slice_k100(w2s, algo):
    _case(f"200k slice k=100 nprobe=16 algo={algo}")
    D, I, st = w2s["idx"].search(300, w2s["xq"], 100, params=hx.make_params(nprobe=16, algo=algo))
    OD, OI, _ = w2s["ix"].search(w2s["xq"], 16, 100)
    assert_same(D, I, OD, OI)
    # FP64 recomputation of every returned distance: integer data, so exact
    ref = ((w2s["xb"][I].astype(np.float64) - w2s["xq"][:, None, :].astype(np.float64)) ** 2).sum(-1)
    assert np.array_equal(ref, D.astype(np.float64))


def test_slice_per_query_nprobe(w2s):
    _case("per-query nprobe")
    npq = ds.per_query_nprobe(300, 1, 40, seed=7)
    p = hx.make_params(algo=hx.ALGO_PLAIN, nprobe_per_query=npq)
    D, I, _ = w2s["idx"].search(300, w2s["xq"], 100, params=p)
    _, probes = oracle.coarse(oracle.METRIC_L2, w2s["xq"], w2s["idx"].centroids, 40)
    for i in range(0, 300, 37):
        OD, OI, _ = w2s["ix"].scan(w2s["xq"][i:i + 1], probes[i:i + 1, :npq[i]], 100)
        assert_same(D[i:i + 1], I[i:i + 1], OD, OI)


# ---------------------------------------------------------------- batched MFMA path
def test_batched_workload1(g1, w1):
    _case("batched (BF16 MFMA, integer data: exact distances) workload 1")
    p = hx.make_params(nprobe=10, algo=hx.ALGO_BATCHED, tau=hx.TAU_WARMUP)
    D, I, st = g1.search(100, w1["xq"], 10, params=p)
    OD, OI, _ = w1["ix"].search(w1["xq"], 10, 10)
    assert_same(D, I, OD, OI)
    assert st.fallback_queries == 0 and st.survivors > 0


@pytest.mark.parametrize("cap", [16, 100000])
def test_batched_slice(w2s, cap):
    _case(f"batched 200k slice k=100 nprobe=16 survivor_cap={cap}")
    p = hx.make_params(nprobe=16, algo=hx.ALGO_BATCHED, survivor_cap=cap)
    D, I, st = w2s["idx"].search(300, w2s["xq"], 100, params=p)
    OD, OI, _ = w2s["ix"].search(w2s["xq"], 16, 100)
    assert_same(D, I, OD, OI)
    if cap == 16:
        assert st.fallback_queries == 300  # every query overflows and is rescanned exactly


def test_batched_per_query_nprobe(w2s):
    _case("batched per-query nprobe")
    npq = ds.per_query_nprobe(300, 1, 40, seed=11)
    D, I, _ = w2s["idx"].search(300, w2s["xq"], 100, params=hx.make_params(
        algo=hx.ALGO_BATCHED, nprobe_per_query=npq))
    P, Pi, _ = w2s["idx"].search(300, w2s["xq"], 100, params=hx.make_params(
        algo=hx.ALGO_PLAIN, nprobe_per_query=npq))
    assert_same(D, I, P, Pi)


@pytest.mark.parametrize("wid,nb,nlist,nprobe", [(3, 20_000, 64, 8), (4, 20_000, 64, 8)])
def test_batched_real_valued(wid, nb, nlist, nprobe):
    _case(f"batched on real-valued workload {wid} recipe (margin + FP32 rescoring)")
    w = ds.WORKLOADS[wid]
    xb, xq = ds.generate(w, nb=nb, nq=64, wid=wid)
    sample, init = ds.kmeans_init(xb, nlist)
    idx = hx.Index(w.d, nlist, nprobe, hx.METRIC_L2)
    idx.train(sample.shape[0], sample, init_centroids=init, niter=3)
    idx.add(nb, xb)
    off, ids, codes = idx.csr()
    ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
    OD, OI, _ = ix.search(xq, nprobe, 100)
    for algo in (hx.ALGO_PLAIN, hx.ALGO_PRUNED, hx.ALGO_BATCHED):
        D, I, st = idx.search(64, xq, 100, params=hx.make_params(nprobe=nprobe, algo=algo,
                                                                  dim_block=w.dim_block))
        assert_same(D, I, OD, OI)
    # every returned distance against an FP64 recomputation: FP32 chain error
    ref = ((xb[I].astype(np.float64) - xq[:, None, :].astype(np.float64)) ** 2).sum(-1)
    rel = np.abs(D - ref) / np.maximum(ref, 1e-30)
    print(f"[fp64] workload {wid}: max relative error of returned distances {rel.max():.3e}")
    # all terms are non-negative: a sequential FP32 chain is within (d+1) u
    assert rel.max() <= (w.d + 1) * 2.0 ** -24 * 1.01


def test_batched_zipf_lists_per_query_nprobe():
    """Workload-5 recipe at 300k rows (Zipf(1.1) list popularity, so list
    lengths differ by orders of magnitude; per-query nprobe in [8, 256]):
    batched warm-up and learned-bound searches equal the plain kernel, and
    the plain kernel equals the oracle on a sample of queries."""
    _case("workload-5 recipe 300k x 128, nlist 512, Zipf lists, per-query nprobe 8..256")
    w = ds.WORKLOADS[5]
    xb, xq = ds.generate(w, nb=300_000, nq=512, wid=5)
    sample, init = ds.kmeans_init(xb, 512)
    idx = hx.Index(128, 512, 64, hx.METRIC_L2)
    idx.train(sample.shape[0], sample, init_centroids=init, niter=4)
    idx.add(xb.shape[0], xb)
    sizes = np.array([idx.list_size(l) for l in range(512)])
    print(f"[zipf] list sizes min {sizes.min()} median {int(np.median(sizes))} max {sizes.max()}", flush=True)
    npq = ds.per_query_nprobe(512, 8, 256, seed=3005)
    P, Pi, _ = idx.search(512, xq, 100, params=hx.make_params(algo=hx.ALGO_PLAIN, nprobe_per_query=npq))
    D, I, st = idx.search(512, xq, 100, params=hx.make_params(algo=hx.ALGO_BATCHED, nprobe_per_query=npq))
    assert_same(D, I, P, Pi)
    _, xt = ds.generate(w, nb=0, nq=4000, wid=5, query_seed=3005)
    idx.train_tau_model(xt, 100, 64, epochs=10)
    D, I, st = idx.search(512, xq, 100, params=hx.make_params(algo=hx.ALGO_BATCHED, tau=hx.TAU_LEARNED,
                                                              nprobe_per_query=npq))
    assert_same(D, I, P, Pi)
    off, ids, codes = idx.csr()
    ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
    _, probes = oracle.coarse(oracle.METRIC_L2, xq, idx.centroids, 256)
    for i in range(0, 512, 61):
        OD, OI, _ = ix.scan(xq[i:i + 1], probes[i:i + 1, :npq[i]], 100)
        assert_same(P[i:i + 1], Pi[i:i + 1], OD, OI)


@pytest.mark.parametrize("d", [128, 200, 256, 7])
def test_batched_int8_cases(d):
    """Integer lists take the int8 scan: d up to 256 (4 or 8 chunks of 32
    dims), values at both ends of [0, 255], and a batch with one non-integer
    query (then the BF16 scan with the error margin runs for every query)."""
    _case(f"batched int8 scan d={d}")
    rng = np.random.default_rng(d)
    nb, nlist = 6000, 24
    xb = rng.integers(0, 256, (nb, d)).astype(np.float32)
    xb[:50] = 0.0
    xb[50:100] = 255.0
    xq = rng.integers(0, 256, (70, d)).astype(np.float32)
    xq[0] = 255.0
    xq[1] = 0.0
    xq[2] = xb[7]
    sample, init = ds.kmeans_init(xb, nlist)
    idx = hx.Index(d, nlist, 6, hx.METRIC_L2)
    idx.train(sample.shape[0], sample, init_centroids=init, niter=3)
    idx.add(nb, xb)
    off, ids, codes = idx.csr()
    ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
    for mixed in (False, True):
        q = xq.copy()
        if mixed:
            q[5, 3] += 0.5  # one query off the integer grid
        OD, OI, _ = ix.search(q, 6, 50)
        for tau in (hx.TAU_WARMUP, hx.TAU_NONE):
            D, I, st = idx.search(70, q, 50, params=hx.make_params(nprobe=6, algo=hx.ALGO_BATCHED, tau=tau))
            assert_same(D, I, OD, OI)


@pytest.mark.parametrize("bq", [32, 64])
def test_batched_int8_query_groups(bq):
    """The int8 scan with one or two 32-query groups per wave (items of 32 or
    64 queries): many queries per list, so both groups of an item are full,
    plus a last partial group."""
    _case(f"batched int8 scan, {bq} queries per item")
    rng = np.random.default_rng(17)
    nb, nlist, d = 8000, 8, 128
    xb = rng.integers(0, 256, (nb, d)).astype(np.float32)
    xq = rng.integers(0, 256, (301, d)).astype(np.float32)
    xq[3] = xb[11]
    sample, init = ds.kmeans_init(xb, nlist)
    idx = hx.Index(d, nlist, 4, hx.METRIC_L2)
    idx.train(sample.shape[0], sample, init_centroids=init, niter=3)
    idx.add(nb, xb)
    off, ids, codes = idx.csr()
    ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
    OD, OI, _ = ix.search(xq, 4, 40)
    for tau in (hx.TAU_WARMUP, hx.TAU_NONE):
        for rows in (256, 2048):
            D, I, st = idx.search(301, xq, 40, params=hx.make_params(
                nprobe=4, algo=hx.ALGO_BATCHED, tau=tau, batch_queries=bq, batch_rows=rows))
            assert st.batch_queries == bq and st.scan_kind == 2
            assert_same(D, I, OD, OI)
    # one query off the integer grid: every query takes the BF16 form with
    # the error margin (run by the int8 launch itself for 64-query items)
    qm = xq.copy()
    qm[7, 5] += 0.25
    OD, OI, _ = ix.search(qm, 4, 40)
    for tau in (hx.TAU_WARMUP, hx.TAU_NONE):
        D, I, st = idx.search(301, qm, 40, params=hx.make_params(
            nprobe=4, algo=hx.ALGO_BATCHED, tau=tau, batch_queries=bq, batch_rows=1024))
        assert st.scan_kind != 2
        assert_same(D, I, OD, OI)


@pytest.mark.parametrize("nlist", [1500, 3000, 5000])
def test_batched_int8_planning_class_counts(nlist):
    """The planning at each number of reach classes: 4 (1024 < nlist <= 2048),
    2 (up to 4096), and the per-list atomic kernels above 4096 lists; the
    batched int8 scan equals the plain kernel, and the oracle on a sample."""
    _case(f"batched int8 planning, nlist {nlist}")
    rng = np.random.default_rng(nlist)
    nb, d, nq, P, k = 30000, 32, 200, 32, 20
    xb = rng.integers(0, 256, (nb, d)).astype(np.float32)
    xq = rng.integers(0, 256, (nq, d)).astype(np.float32)
    sample, init = ds.kmeans_init(xb, nlist)
    idx = hx.Index(d, nlist, P, hx.METRIC_L2)
    idx.train(sample.shape[0], sample, init_centroids=init, niter=2)
    idx.add(nb, xb)
    PD, PI, _ = idx.search(nq, xq, k, params=hx.make_params(algo=hx.ALGO_PLAIN, nprobe=P))
    off, ids, codes = idx.csr()
    ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
    OD, OI, _ = ix.search(xq[:25], P, k)
    assert_same(PD[:25], PI[:25], OD, OI)
    for tau in (hx.TAU_WARMUP, hx.TAU_NONE):
        D, I, st = idx.search(nq, xq, k, params=hx.make_params(nprobe=P, algo=hx.ALGO_BATCHED, tau=tau))
        assert st.scan_kind == 2
        assert_same(D, I, PD, PI)


def test_batched_int8_duplicate_rows():
    """Every row repeated about 60 times: equal radii within a list (the int8
    shadow's radius order falls back to arena order) and runs of equal
    distances across the k-th position, which the (distance, id) order
    settles; the batched int8 scan equals the oracle at both item sizes."""
    _case("batched int8 scan, duplicated rows")
    rng = np.random.default_rng(5)
    d, nlist, P, k = 64, 16, 6, 90
    base = rng.integers(0, 256, (150, d)).astype(np.float32)
    xb = base[rng.integers(0, 150, 9000)]
    xq = np.concatenate([base[:40], rng.integers(0, 256, (40, d)).astype(np.float32)])
    sample, init = ds.kmeans_init(xb, nlist)
    idx = hx.Index(d, nlist, P, hx.METRIC_L2)
    idx.train(sample.shape[0], sample, init_centroids=init, niter=3)
    idx.add(xb.shape[0], xb)
    off, ids, codes = idx.csr()
    ix = oracle.Index(oracle.METRIC_L2, idx.centroids, off, ids, codes)
    OD, OI, _ = ix.search(xq, P, k)
    assert (OD[:40, 0] == 0.0).all()  # each of the first queries is a stored row
    for tau in (hx.TAU_WARMUP, hx.TAU_NONE):
        for bq in (32, 64):
            D, I, st = idx.search(xq.shape[0], xq, k, params=hx.make_params(
                nprobe=P, algo=hx.ALGO_BATCHED, tau=tau, batch_queries=bq))
            assert st.scan_kind == 2
            assert_same(D, I, OD, OI)
