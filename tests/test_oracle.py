"""Pins for the CPU oracle (oracle/).  CPU only.

Each oracle function is checked against something independent of its own
formula: exact integer arithmetic, FP64 brute force in numpy, the
reference's own heap (oracle/_ref, built from src/heap.hpp), the golden
vectors in tests/golden, and the reference's property test
(src/tests/heap_test.cpp:32-47).  test_planted_mistakes shows that every
planted mistake in the oracle fails at least one of these pins.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, brute_topk, random_partials
from harmonyclient_amd import datasets as ds
from harmonyclient_amd import fvecs

L2, IP = oracle.METRIC_L2, oracle.METRIC_IP
FMAX = np.finfo(np.float32).max


# ---------------------------------------------------------------- distances
def pin_distance_integer(seed=0):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (50, 128)).astype(np.float32)
    b = rng.integers(0, 256, (50, 128)).astype(np.float32)
    ok = True
    for i in range(50):
        exact = int(((a[i].astype(np.int64) - b[i].astype(np.int64)) ** 2).sum())
        ok &= oracle.distance(L2, a[i], b[i]) == exact
        ok &= oracle.distance(IP, a[i], b[i]) == int((a[i].astype(np.int64) * b[i].astype(np.int64)).sum())
    return ok


def test_distance_exact_on_integers():
    # values 0..255, d = 128: every partial sum < 2^24, so FP32 is exact
    assert pin_distance_integer()


def test_distance_error_bound_on_reals():
    rng = np.random.default_rng(1)
    for d in (3, 70, 420):
        a = rng.standard_t(3, (20, d)).astype(np.float32)
        b = rng.standard_t(3, (20, d)).astype(np.float32)
        for i in range(20):
            t = a[i].astype(np.float64) - b[i].astype(np.float64)
            exact = (t * t).sum()
            # sequential FMA chain: |err| <= (d + 1) u sum|terms| (Higham 3.5)
            bound = (d + 1) * 2.0 ** -24 * exact * 1.01 + 1e-30
            assert abs(oracle.distance(L2, a[i], b[i]) - exact) <= bound
            ip = (a[i].astype(np.float64) * b[i].astype(np.float64))
            assert abs(oracle.distance(IP, a[i], b[i]) - ip.sum()) <= (d + 1) * 2.0 ** -24 * np.abs(ip).sum() * 1.01


# ---------------------------------------------------------------- top-k heap
def pin_topk_vs_sort(seed=0):
    rng = np.random.default_rng(seed)
    ok = True
    for metric in (L2, IP):
        for k in (1, 7, 20, 64):
            vals = rng.integers(0, 30, 300).astype(np.float32)  # many ties
            ids = rng.permutation(300).astype(np.int64)
            ov, oi = oracle.topk(metric, vals, ids, k)
            key = vals if metric == L2 else -vals
            order = np.lexsort((ids, key))[:k]
            ok &= np.array_equal(ov, vals[order]) and np.array_equal(oi, ids[order])
    return ok


def test_topk_is_lexicographic_sort():
    assert pin_topk_vs_sort()


def test_topk_sentinels():
    ov, oi = oracle.topk(L2, np.array([3.0, 1.0], np.float32), np.array([5, 6]), 4)
    assert ov.tolist() == [1.0, 3.0, FMAX, FMAX] and oi.tolist() == [6, 5, -1, -1]
    ov, oi = oracle.topk(IP, np.array([3.0, 1.0], np.float32), np.array([5, 6]), 3)
    assert ov.tolist() == [3.0, 1.0, -FMAX] and oi.tolist() == [5, 6, -1]
    # inf and NaN are never admitted (strict admission against FLT_MAX)
    ov, oi = oracle.topk(L2, np.array([np.inf, np.nan, FMAX, 2.0], np.float32), np.arange(4), 3)
    assert oi.tolist() == [3, -1, -1]


def test_heap_property_like_reference():
    """src/tests/heap_test.cpp:7-48: top-k of n random values equals the
    first k of the sorted values, and each value matches its id."""
    rng = np.random.default_rng(2)
    for metric in (L2, IP):
        vals = rng.random(100).astype(np.float32)
        ov, oi = oracle.topk(metric, vals, np.arange(100), 20)
        assert np.array_equal(vals[oi], ov)
        srt = np.sort(vals) if metric == L2 else np.sort(vals)[::-1]
        assert np.array_equal(ov, srt[:20])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref needs the reference tree")
def test_topk_vs_reference_heap():
    """The reference's own heap (src/heap.hpp): same distances always, same
    ids whenever no two values are equal."""
    rng = np.random.default_rng(3)
    for metric in (L2, IP):
        for k in (1, 10, 100):
            vals = rng.random(1000).astype(np.float32)
            ids = np.arange(1000)
            a = oracle.topk(metric, vals, ids, k)
            b = oracle.ref_topk(metric, vals, ids, k)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            tv = rng.integers(0, 20, 1000).astype(np.float32)
            a = oracle.topk(metric, tv, ids, k)
            b = oracle.ref_topk(metric, tv, ids, k)
            assert np.array_equal(a[0], b[0])


# ---------------------------------------------------------------- top-k merge
def sort_merge(metric, Dp, Ip, k):
    """Merge by a full lexicographic sort of the valid entries (numpy)."""
    parts, nq, k_in = Dp.shape
    D = np.full((nq, k), FMAX if metric == L2 else -FMAX, np.float32)
    I = np.full((nq, k), -1, np.int64)
    for q in range(nq):
        v, i = Dp[:, q].ravel(), Ip[:, q].ravel()
        keep = i >= 0
        v, i = v[keep], i[keep]
        o = np.lexsort((i, v if metric == L2 else -v))[:k]
        D[q, :o.size], I[q, :o.size] = v[o], i[o]
    return D, I


def pin_merge_vs_sort():
    ok = True
    for metric in (L2, IP):
        for parts, k_in, k in ((1, 10, 10), (4, 30, 20), (7, 16, 64), (3, 50, 1)):
            Dp, Ip = random_partials(parts * 100 + k, parts, 25, k_in, metric)
            D, I = oracle.merge(metric, Dp, Ip, k)
            SD, SI = sort_merge(metric, Dp, Ip, k)
            ok &= np.array_equal(D, SD) and np.array_equal(I, SI)
    return ok


def shard_index(ix, owner, s):
    """The part of a CSR oracle index whose lists satisfy owner[l] == s (the
    other lists are empty); centroids are shared."""
    off, ids, codes = ix.list_off, ix.ids, ix.codes
    sizes = np.diff(off) * (owner == s)
    noff = np.concatenate([[0], np.cumsum(sizes)])
    rows = np.concatenate([np.arange(off[l], off[l + 1]) for l in range(ix.nlist) if owner[l] == s] or
                          [np.zeros(0, np.int64)])
    return oracle.Index(ix.metric, ix.centroids, noff, ids[rows], codes[rows])


def pin_merge_sharding(ix, q, nprobe, k, shards=3):
    """List-parallel sharding identity: searching each shard of the lists
    with the same probes and merging equals searching all lists."""
    _, probes = oracle.coarse(ix.metric, q, ix.centroids, nprobe)
    owner = np.arange(ix.nlist) % shards
    parts = [shard_index(ix, owner, s).scan(q, probes, k) for s in range(shards)]
    D, I = oracle.merge(ix.metric, np.stack([p[0] for p in parts]), np.stack([p[1] for p in parts]), k)
    FD, FI, _ = ix.scan(q, probes, k)
    return np.array_equal(D, FD) and np.array_equal(I, FI)


def test_merge_vs_sort():
    assert pin_merge_vs_sort()


def test_merge_sharding_identity(w1):
    assert pin_merge_sharding(w1["ix"], w1["xq"], 10, 10)
    assert pin_merge_sharding(w1["ix"], w1["xq"][:20], 40, 100, shards=5)


def test_merge_edges():
    # all partitions empty -> sentinels; k larger than the valid entries
    Dp = np.full((3, 2, 4), FMAX, np.float32)
    Ip = np.full((3, 2, 4), -1, np.int64)
    Dp[1, 0, 0], Ip[1, 0, 0] = 5.0, 7
    D, I = oracle.merge(L2, Dp, Ip, 3)
    assert I.tolist() == [[7, -1, -1], [-1, -1, -1]] and D[0].tolist() == [5.0, FMAX, FMAX]


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref needs the reference tree")
def test_merge_vs_reference_heap():
    """The reference merges with its own heap (src/Index.cpp:1327-1343): the
    entries of each partition offered in order with strict admission.  Same
    distances; same ids when no two values are equal."""
    for metric in (L2, IP):
        for ties in (False, True):
            Dp, Ip = random_partials(9, 5, 20, 40, metric, ties=ties)
            D, I = oracle.merge(metric, Dp, Ip, 30)
            for q in range(20):
                rd, ri = oracle.ref_topk(metric, Dp[:, q].ravel(), Ip[:, q].ravel(), 30)
                assert np.array_equal(D[q], rd)
                if not ties:
                    assert np.array_equal(I[q], ri)


# ---------------------------------------------------------------- coarse / build
def pin_coarse(w1_small):
    xq, cent = w1_small
    od, ol = oracle.coarse(L2, xq, cent, 10)
    ok = True
    for i in range(xq.shape[0]):
        dist, ids = brute_topk(cent, xq[i], 10)
        ok &= np.array_equal(ol[i], ids) and np.array_equal(od[i].astype(np.float64), dist)
    return ok


@pytest.fixture(scope="module")
def int_centroids(w1):
    # integer-valued centroids keep every distance exact, so FP64 brute force
    # pins ids including ties
    return np.rint(w1["cent"]).astype(np.float32)


def test_coarse_vs_bruteforce(w1, int_centroids):
    assert pin_coarse((w1["xq"], int_centroids))


def test_coarse_clamps_nprobe(w1):
    od, ol = oracle.coarse(L2, w1["xq"][:3], w1["cent"], 5000)
    assert ol.shape == (3, 100) and np.array_equal(np.sort(ol[0]), np.arange(100))


def test_build_lists(w1, int_centroids):
    xb = w1["xb"]
    off, order = oracle.build_lists(L2, xb, int_centroids)
    X = xb.astype(np.float64)
    C = int_centroids.astype(np.float64)
    d2 = (X * X).sum(1)[:, None] + (C * C).sum(1)[None, :] - 2 * X @ C.T
    assign = d2.argmin(1)  # argmin returns the lowest id among equal minima
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(assign, minlength=100))]))
    for l in range(100):
        rows = order[off[l]:off[l + 1]]
        assert np.all(assign[rows] == l)
        key = np.lexsort((rows, d2[rows, l]))
        assert np.array_equal(key, np.arange(rows.size))  # sorted by (distance, row)


def test_kmeans_one_round_is_lloyd(w1):
    s, init = w1["sample"], w1["init"]
    got = oracle.kmeans(L2, s, init, 1)
    _, assign = oracle.assign(L2, s, init)
    want = init.copy()
    for c in range(init.shape[0]):
        m = assign == c
        if m.any():
            want[c] = (s[m].astype(np.float64).sum(0) / m.sum()).astype(np.float32)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------- scans
def union_rows(ix, probes):
    return np.concatenate([np.arange(ix.list_off[l], ix.list_off[l + 1]) for l in probes if l >= 0])


def pin_search_vs_bruteforce(ix, xq, nprobe, k):
    D, I, st = ix.search(xq, nprobe, k)
    _, probes = oracle.coarse(ix.metric, xq, ix.centroids, nprobe)
    ok = True
    for i in range(xq.shape[0]):
        rows = union_rows(ix, probes[i])
        dist, ids = brute_topk(ix.codes, xq[i], k, metric=ix.metric, rows=rows, ids=ix.ids[rows])
        n = min(k, rows.size)
        ok &= np.array_equal(I[i, :n], ids) and np.array_equal(D[i, :n].astype(np.float64), dist)
        ok &= bool(np.all(I[i, n:] == -1))
    return ok


def test_search_vs_bruteforce(w1):
    assert pin_search_vs_bruteforce(w1["ix"], w1["xq"], 10, 10)
    assert pin_search_vs_bruteforce(w1["ix"], w1["xq"][:20], 3, 100)


def test_search_ip_vs_bruteforce(w1):
    cent = oracle.kmeans(IP, w1["sample"], w1["init"], 3)
    ix = oracle.Index.build(IP, w1["xb"], cent)
    assert pin_search_vs_bruteforce(ix, w1["xq"][:30], 5, 10)


def test_pruned_equals_plain(w1):
    ix = w1["ix"]
    _, probes = oracle.coarse(L2, w1["xq"], w1["cent"], 10)
    D, I, _ = ix.scan(w1["xq"], probes, 10)
    tau = ix.warmup_tau(w1["xq"], probes, 10, 10, 10)
    for running in (True, False):
        for blk in (8, 32, 50, 128):
            for t in (tau, None):
                PD, PI, st = ix.search_pruned(w1["xq"], probes, 10, blk, t, running)
                assert np.array_equal(PD, D) and np.array_equal(PI, I)
                if blk < 128 and (running or t is not None):
                    assert st.blocks_skipped > 0


def pin_prune_ties():
    """Row 0 (scanned second) ties row 1 at distance 1 and its running
    distance equals the bound after the first block: it must not be pruned,
    and it wins the tie by id (src/node.cpp:647 prunes only when strictly
    greater)."""
    q = np.zeros((1, 8), np.float32)
    codes = np.zeros((2, 8), np.float32)
    codes[0, 0] = 1.0  # id 0, partial after dims 0..3 = 1 = final
    codes[1, 1] = 1.0  # id 1, same distance
    ix = oracle.Index(L2, np.zeros((1, 8), np.float32), [0, 2], [1, 0], codes[[1, 0]])
    D, I, _ = ix.search_pruned(q, np.zeros((1, 1), np.int64), 1, 4, np.array([1.0], np.float32), True)
    return I[0, 0] == 0 and D[0, 0] == 1.0


def test_prune_is_strict():
    assert pin_prune_ties()


def pin_warmup(ix, xq, probes, k, L, W):
    tau = ix.warmup_tau(xq, probes, k, L, W)
    ok = True
    for i in range(xq.shape[0]):
        rows = np.concatenate([np.arange(ix.list_off[l], ix.list_off[l] + min(W, ix.list_off[l + 1] - ix.list_off[l]))
                               for l in probes[i][:L]])
        dist = ((ix.codes[rows].astype(np.float64) - xq[i].astype(np.float64)) ** 2).sum(1)
        want = np.sort(dist)[k - 1] if rows.size >= k else FMAX
        ok &= float(tau[i]) == float(np.float32(want))
    return ok


def test_warmup_tau_vs_bruteforce(w1):
    _, probes = oracle.coarse(L2, w1["xq"], w1["cent"], 10)
    for L, W in [(10, 10), (3, 50), (1, 5), (20, 1000)]:
        assert pin_warmup(w1["ix"], w1["xq"], probes, 10, L, W)


def test_warmup_tau_bounds_kth(w1):
    ix = w1["ix"]
    _, probes = oracle.coarse(L2, w1["xq"], w1["cent"], 10)
    D, _, _ = ix.scan(w1["xq"], probes, 10)
    tau = ix.warmup_tau(w1["xq"], probes, 10, 10, 10)
    assert np.all(tau >= D[:, -1])


# ---------------------------------------------------------------- golden / reference
def test_golden_workload1(w1):
    g = np.load(os.path.join(GOLDEN, "workload1_golden.npz"))
    assert np.array_equal(g["centroids"], w1["cent"])
    D, I, _ = w1["ix"].search(w1["xq"], 10, 10)
    assert np.array_equal(I, g["I"]) and np.array_equal(D, g["D"])
    _, probes = oracle.coarse(L2, w1["xq"], w1["cent"], 10)
    assert np.array_equal(probes, g["probes"])
    assert np.array_equal(w1["ix"].warmup_tau(w1["xq"], probes, 10, 10, 10), g["tau"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref needs the reference tree")
def test_search_vs_reference_build(w1):
    """The reference's plain search (its heap, its scan order): identical
    distances; the same (distance, id) pairs strictly inside the k-th
    distance (the reference orders equal distances by heap internals, and
    at the k-th it keeps whichever tied rows its heap retained)."""
    xb, xq = w1["xb"], w1["xq"]
    ix = w1["ix"]
    for nprobe, k in [(10, 10), (5, 100), (100, 50)]:
        D, I, _ = ix.search(xq, nprobe, k)
        RD, RI, RP = ix.ref_search(xq, nprobe, k)
        assert np.array_equal(D, RD)
        for i in range(xq.shape[0]):
            strict = D[i] < D[i, -1]
            ours = sorted(zip(D[i, strict].tolist(), I[i, strict].tolist()))
            theirs = sorted(zip(RD[i, strict].tolist(), RI[i, strict].tolist()))
            assert ours == theirs
        _, probes = oracle.coarse(L2, xq, w1["cent"], nprobe)
        assert np.array_equal(RP, probes)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_cpu_baselines_base1_and_faiss_algorithm(w1):
    """The two other CPU baselines in oracle/_ref return the plain search's
    distances: Harmony base with one worker (warm-up bound + four-check
    pruning, reference heap) and the faiss IndexIVFFlat algorithm (GEMM-form
    coarse quantizer, (distance, id) heap -- integer data, so its coarse
    distances and probe sets are exact and its ids equal the oracle's)."""
    xq = w1["xq"]
    ix = w1["ix"]
    for nprobe, k in [(10, 10), (5, 100), (100, 50)]:
        D, I, _ = ix.search(xq, nprobe, k)
        BD, BI = ix.ref_search_base1(xq, nprobe, k)
        assert np.array_equal(D, BD)
        for i in range(xq.shape[0]):
            strict = D[i] < D[i, -1]
            assert sorted(zip(D[i, strict].tolist(), I[i, strict].tolist())) == \
                sorted(zip(BD[i, strict].tolist(), BI[i, strict].tolist()))
        FD, FI = ix.ref_search_faiss(xq, nprobe, k)
        assert np.array_equal(D, FD) and np.array_equal(I, FI)


def test_reference_fixtures_sift10k(w1):
    xq = fvecs.read_fvecs(os.path.join(GOLDEN, "reference_fixtures", "sift10k_query.fvecs"))
    assert xq.shape == (100, 128) and np.all(xq == np.rint(xq))
    assert pin_search_vs_bruteforce(w1["ix"], xq, 10, 10)


def test_reference_fixtures_small():
    root = os.path.join(GOLDEN, "reference_fixtures")
    iris = fvecs.read_fvecs(os.path.join(root, "iris_base.fvecs"))
    assert iris.shape == (150, 4)
    t = fvecs.read_fvecs(os.path.join(root, "test_base.fvecs"), leading_word=True)
    assert t.shape == (5, 20)
    # iris holds one decimal digit: x10 makes it integer-valued, so FP32
    # distances are exact and the FP64 brute force pins ids and ties
    iris10 = np.rint(iris * 10).astype(np.float32)
    q10 = np.rint(fvecs.read_fvecs(os.path.join(root, "iris_query.fvecs")) * 10).astype(np.float32)
    ix = oracle.Index.build(L2, iris10, iris10[[0, 50, 100]])
    assert pin_search_vs_bruteforce(ix, q10, 2, 5)
    ix = oracle.Index.build(L2, t, t[[0, 3]])
    assert pin_search_vs_bruteforce(ix, t, 1, 3)


def test_recall_definition():
    I = np.array([[1, 2, 3], [4, -1, 9]])
    GT = np.array([[3, 2, 7], [4, 5, 6]])
    assert oracle.recall(I, GT) == pytest.approx(3 / 6)


# ---------------------------------------------------------------- planted mistakes
def all_pins(w1, int_centroids):
    _, probes = oracle.coarse(L2, w1["xq"], w1["cent"], 10)
    sub = w1["xq"][:30]
    return {
        "distance": pin_distance_integer(),
        "topk": pin_topk_vs_sort(),
        "coarse": pin_coarse((w1["xq"][:30], int_centroids)),
        "search": pin_search_vs_bruteforce(w1["ix"], sub, 10, 10),
        "prune_ties": pin_prune_ties(),
        "warmup": pin_warmup(w1["ix"], sub, probes[:30], 10, 10, 10),
        "merge": pin_merge_vs_sort(),
        "shard": pin_merge_sharding(w1["ix"], sub, 10, 10),
    }


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes(w1, int_centroids, m):
    assert all(all_pins(w1, int_centroids).values())
    with oracle.mutation(m):
        pins = all_pins(w1, int_centroids)
    assert not all(pins.values()), f"mutation {m} went unnoticed"
