"""Seeded synthetic datasets for the IVF-Flat workloads.

Every generator is deterministic in its seeds (numpy PCG64), so the same
arrays are produced on any host.  Seeds: base = 1000 + workload number,
queries = 2000 + workload number, k-means = 6666 (the reference's
``cp.seed``, src/Index.cpp:393).

Workloads (base x dim, lists, probes, k):
  1  10,000 x 128, integer-valued 0..255 Gaussian mixture (100 centres),
     100 queries, nlist 100, nprobe 10, k 10, dimension blocks of 32.
  2  1,000,000 x 128, same recipe with 1024 centres, 10,000 queries,
     nlist 1024, nprobe in {16, 64, 128}, k 100, 4 blocks of 32.
  3  1,000,000 x 420 heavy-tailed reals (Student-t nu=3 per dim, log-normal
     per-dim scale, 1000 centres), nlist 1000, nprobe 100, k 100.
  4  270,000 x 500 non-negative 70 % sparse, L2-normalised rows (500 centres),
     nlist 512, nprobe 64, k 100.
  5  10,000,000 x 128 integer mixture with Zipf(1.1) centre popularity,
     nlist 4096, per-query nprobe in [8, 256], k 100.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Workload:
    name: str
    nb: int
    d: int
    nq: int
    ncomp: int
    nlist: int
    nprobe: int
    k: int
    dim_block: int
    kind: str  # "int_mixture", "student_t", "sparse_nonneg"
    zipf: float = 0.0
    kmeans_iters: int = 10


WORKLOADS = {
    1: Workload("ivf_flat_10k_d128", 10_000, 128, 100, 100, 100, 10, 10, 32, "int_mixture"),
    2: Workload("ivf_flat_1m_d128", 1_000_000, 128, 10_000, 1024, 1024, 64, 100, 32,
                "int_mixture"),
    3: Workload("ivf_flat_1m_d420", 1_000_000, 420, 10_000, 1000, 1000, 100, 100, 70,
                "student_t"),
    4: Workload("ivf_flat_270k_d500", 270_000, 500, 20_000, 500, 512, 64, 100, 125,
                "sparse_nonneg"),
    5: Workload("ivf_flat_10m_d128_zipf", 10_000_000, 128, 100_000, 4096, 4096, 0, 100, 32,
                "int_mixture", zipf=1.1),
}


def _int_mixture(rng, centres, n, sigma=20.0, weights=None, chunk=1 << 18):
    d = centres.shape[1]
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        if weights is None:
            comp = rng.integers(0, centres.shape[0], e - s)
        else:
            comp = rng.choice(centres.shape[0], size=e - s, p=weights)
        x = centres[comp] + rng.normal(0.0, sigma, (e - s, d))
        out[s:e] = np.clip(np.rint(x), 0, 255)
    return out


def _student_t(rng, centres, scale, n, chunk=1 << 17):
    d = centres.shape[1]
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = rng.integers(0, centres.shape[0], e - s)
        out[s:e] = centres[comp] + rng.standard_t(3.0, (e - s, d)) * scale
    return out


def _sparse_nonneg(rng, centres, n, chunk=1 << 16):
    d = centres.shape[1]
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = rng.integers(0, centres.shape[0], e - s)
        mask = rng.random((e - s, d)) < centres[comp]
        x = np.where(mask, rng.exponential(1.0, (e - s, d)), 0.0)
        nrm = np.sqrt((x * x).sum(axis=1, keepdims=True))
        nrm[nrm == 0] = 1.0
        out[s:e] = x / nrm
    return out


def generate(w: Workload | int, nb: int | None = None, nq: int | None = None, wid: int | None = None,
             query_seed: int | None = None):
    """Return (base [nb x d] float32, queries [nq x d] float32).

    nb / nq override the workload's sizes (same recipe, fewer rows).
    query_seed replaces the query stream's seed (2000 + workload number),
    e.g. 3000 + workload number for the bound model's training queries."""
    if isinstance(w, int):
        wid = w
        w = WORKLOADS[w]
    wid = wid if wid is not None else next(i for i, v in WORKLOADS.items() if v is w)
    nb = w.nb if nb is None else nb
    nq = w.nq if nq is None else nq
    rb = np.random.default_rng(1000 + wid)
    rq = np.random.default_rng(2000 + wid if query_seed is None else query_seed)
    if w.kind == "int_mixture":
        centres = rb.uniform(0.0, 255.0, (w.ncomp, w.d))
        weights = None
        if w.zipf > 0:
            p = 1.0 / np.arange(1, w.ncomp + 1) ** w.zipf
            weights = p / p.sum()
        return _int_mixture(rb, centres, nb, weights=weights), _int_mixture(rq, centres, nq,
                                                                            weights=weights)
    if w.kind == "student_t":
        centres = rb.normal(0.0, 3.0, (w.ncomp, w.d))
        scale = np.exp(rb.normal(0.0, 0.5, w.d))
        return _student_t(rb, centres, scale, nb), _student_t(rq, centres, scale, nq)
    if w.kind == "sparse_nonneg":
        # per-centre probability of a nonzero, mean 0.3 (70 % sparse)
        centres = np.clip(rb.beta(0.6, 1.4, (w.ncomp, w.d)), 0.0, 1.0)
        return _sparse_nonneg(rb, centres, nb), _sparse_nonneg(rq, centres, nq)
    raise ValueError(w.kind)


def per_query_nprobe(nq: int, lo: int = 8, hi: int = 256, seed: int = 3005) -> np.ndarray:
    """Per-query probe counts uniform in [lo, hi] (workload 5)."""
    return np.random.default_rng(seed).integers(lo, hi + 1, nq).astype(np.int32)


def kmeans_init(x: np.ndarray, nlist: int, max_points_per_centroid: int = 256,
                seed: int = 6666):
    """Training sample and initial centroids, following the reference's
    Clustering::train: subsample to at most nlist * 256 rows
    (src/Clustering.cpp:48-68), then take nlist distinct sample rows as the
    initial centroids (src/Clustering.cpp:72-89).  The permutation is
    numpy's, seeded, not libstdc++'s std::shuffle."""
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    cap = nlist * max_points_per_centroid
    sample = x
    if n > cap:
        idx = rng.permutation(n)[:cap]
        sample = x[np.sort(idx)]
    init = sample[rng.permutation(sample.shape[0])[:nlist]].copy()
    return np.ascontiguousarray(sample), np.ascontiguousarray(init)
