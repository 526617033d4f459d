"""CPU oracle for the IVF-Flat query path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, as the checker.  The product path
(``harmonyclient_amd``) never imports it.

``liboracle.so`` is the C restatement in ``oracle.c`` (reference citations in
``oracle.h``).  ``_ref/libref.so`` is the reference's own ``src/heap.hpp``
compiled where it lies, driven by ``ref_driver.cpp``; it exists only where
it was built (see ``Makefile``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

METRIC_IP = 0
METRIC_L2 = 1

_f32p = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)


class OrcIndex(C.Structure):
    _fields_ = [
        ("metric", C.c_int),
        ("d", C.c_int64),
        ("nlist", C.c_int64),
        ("centroids", _f32p),
        ("list_off", _i64p),
        ("ids", _i64p),
        ("codes", _f32p),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("candidates", C.c_int64),
        ("blocks_total", C.c_int64),
        ("blocks_skipped", C.c_int64),
    ]


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref/libref.so where the reference exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _fp(a):
    return a.ctypes.data_as(_f32p)


def _ip(a):
    return a.ctypes.data_as(_i64p)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_get_mutation.restype = C.c_int
        L.orc_distance.restype = C.c_float
        L.orc_distance.argtypes = [C.c_int, _f32p, _f32p, C.c_int64]
        L.orc_topk.argtypes = [C.c_int, C.c_int64, _f32p, _i64p, C.c_int64, _f32p, _i64p]
        L.orc_coarse.restype = C.c_int64
        L.orc_coarse.argtypes = [C.c_int, C.c_int64, _f32p, C.c_int64, _f32p, C.c_int64,
                                 C.c_int64, _f32p, _i64p]
        L.orc_assign.argtypes = [C.c_int, C.c_int64, _f32p, C.c_int64, _f32p, C.c_int64,
                                 _f32p, _i64p]
        L.orc_build_lists.argtypes = [C.c_int, C.c_int64, _f32p, C.c_int64, _f32p, C.c_int64,
                                      _i64p, _i64p]
        L.orc_warmup_tau.argtypes = [C.POINTER(OrcIndex), C.c_int64, _f32p, _i64p, C.c_int64,
                                     C.c_int64, C.c_int64, C.c_int64, _f32p]
        L.orc_search.restype = C.c_int64
        L.orc_search.argtypes = [C.POINTER(OrcIndex), C.c_int64, _f32p, C.c_int64, C.c_int64,
                                 _f32p, _i64p, C.c_int, C.POINTER(OrcStats)]
        L.orc_search_pruned.argtypes = [C.POINTER(OrcIndex), C.c_int64, _f32p, _i64p, C.c_int64,
                                        C.c_int64, C.c_int64, _f32p, C.c_int, _f32p, _i64p,
                                        C.c_int, C.POINTER(OrcStats)]
        L.orc_scan.argtypes = [C.POINTER(OrcIndex), C.c_int64, _f32p, _i64p, C.c_int64,
                               C.c_int64, _f32p, _i64p, C.c_int, C.POINTER(OrcStats)]
        L.orc_kmeans.argtypes = [C.c_int, C.c_int64, _f32p, C.c_int64, C.c_int64, C.c_int, _f32p]
        L.orc_merge.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int64, _f32p, _i64p, C.c_int64,
                                _f32p, _i64p]
        L.orc_recall.restype = C.c_double
        L.orc_recall.argtypes = [_i64p, _i64p, C.c_int64, C.c_int64, C.c_int64]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_topk.argtypes = [C.c_int, C.c_int64, _f32p, _i64p, C.c_int64, _f32p, _i64p]
        R.ref_search.restype = C.c_int64
        R.ref_search.argtypes = [C.c_int, C.c_int64, _f32p, C.c_int64, C.c_int64, _f32p, _i64p,
                                 _i64p, _f32p, C.c_int64, C.c_int64, _f32p, _i64p, _i64p, C.c_int]
        R.ref_search_base1.restype = C.c_int64
        R.ref_search_base1.argtypes = [C.c_int64, _f32p, C.c_int64, C.c_int64, _f32p, _i64p, _i64p, _f32p,
                                       C.c_int64, C.c_int64, C.c_int64, C.c_int64, _f32p, _i64p, C.c_int]
        R.ref_search_faiss.restype = C.c_int64
        R.ref_search_faiss.argtypes = [C.c_int64, _f32p, C.c_int64, C.c_int64, _f32p, _i64p, _i64p, _f32p,
                                       C.c_int64, C.c_int64, _f32p, _i64p, C.c_int]
        _ref = R
    return _ref


class mutation:
    """Context manager planting a mistake in the oracle (tests only)."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        lib().orc_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().orc_set_mutation(0)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def distance(metric, a, b):
    a, b = f32(a), f32(b)
    return lib().orc_distance(metric, _fp(a), _fp(b), a.shape[-1])


def topk(metric, vals, ids, k):
    vals, ids = f32(vals), i64(ids)
    ov = np.empty(k, np.float32)
    oi = np.empty(k, np.int64)
    lib().orc_topk(metric, vals.size, _fp(vals), _ip(ids), k, _fp(ov), _ip(oi))
    return ov, oi


def merge(metric, D_parts, I_parts, k):
    """Merge [parts x nq x k_in] partial results into [nq x k]."""
    Dp, Ip = f32(D_parts), i64(I_parts)
    parts, nq, k_in = Dp.shape
    D = np.empty((nq, k), np.float32)
    I = np.empty((nq, k), np.int64)
    lib().orc_merge(metric, nq, parts, k_in, _fp(Dp), _ip(Ip), k, _fp(D), _ip(I))
    return D, I


def ref_topk(metric, vals, ids, k):
    vals, ids = f32(vals), i64(ids)
    ov = np.empty(k, np.float32)
    oi = np.empty(k, np.int64)
    ref().ref_topk(metric, vals.size, _fp(vals), _ip(ids), k, _fp(ov), _ip(oi))
    return ov, oi


def coarse(metric, q, cent, nprobe):
    q, cent = f32(q), f32(cent)
    nq, d = q.shape
    nlist = cent.shape[0]
    P = nlist if (nprobe <= 0 or nprobe > nlist) else nprobe
    od = np.empty((nq, P), np.float32)
    ol = np.empty((nq, P), np.int64)
    lib().orc_coarse(metric, nq, _fp(q), nlist, _fp(cent), d, nprobe, _fp(od), _ip(ol))
    return od, ol


def assign(metric, x, cent):
    x, cent = f32(x), f32(cent)
    od = np.empty(x.shape[0], np.float32)
    ol = np.empty(x.shape[0], np.int64)
    lib().orc_assign(metric, x.shape[0], _fp(x), cent.shape[0], _fp(cent), x.shape[1],
                     _fp(od), _ip(ol))
    return od, ol


def build_lists(metric, x, cent):
    x, cent = f32(x), f32(cent)
    nlist = cent.shape[0]
    off = np.empty(nlist + 1, np.int64)
    order = np.empty(x.shape[0], np.int64)
    lib().orc_build_lists(metric, x.shape[0], _fp(x), nlist, _fp(cent), x.shape[1],
                          _ip(off), _ip(order))
    return off, order


def kmeans(metric, x, init_centroids, niter):
    x = f32(x)
    cent = f32(init_centroids).copy()
    lib().orc_kmeans(metric, x.shape[0], _fp(x), cent.shape[0], x.shape[1], niter, _fp(cent))
    return cent


def recall(I, GT, k=None):
    I, GT = i64(I), i64(GT)
    nq = I.shape[0]
    k = k or I.shape[1]
    return lib().orc_recall(_ip(I), _ip(GT), nq, k, GT.shape[1])


class Index:
    """CSR index in the oracle's layout: list l = rows list_off[l]:list_off[l+1]."""

    def __init__(self, metric, centroids, list_off, ids, codes):
        self.metric = metric
        self.centroids = f32(centroids)
        self.list_off = i64(list_off)
        self.ids = i64(ids)
        self.codes = f32(codes)
        self.d = self.centroids.shape[1]
        self.nlist = self.centroids.shape[0]
        self._s = OrcIndex(metric, self.d, self.nlist, _fp(self.centroids), _ip(self.list_off),
                           _ip(self.ids), _fp(self.codes))

    @classmethod
    def build(cls, metric, x, centroids):
        off, order = build_lists(metric, x, centroids)
        x = f32(x)
        return cls(metric, centroids, off, order, x[order])

    def search(self, q, nprobe, k, threads=0):
        q = f32(q)
        nq = q.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        st = OrcStats()
        lib().orc_search(C.byref(self._s), nq, _fp(q), nprobe, k, _fp(D), _ip(I), threads,
                         C.byref(st))
        return D, I, st

    def scan(self, q, probes, k, threads=0):
        q, probes = f32(q), i64(probes)
        nq, P = probes.shape
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        st = OrcStats()
        lib().orc_scan(C.byref(self._s), nq, _fp(q), _ip(probes), P, k, _fp(D), _ip(I),
                       threads, C.byref(st))
        return D, I, st

    def search_pruned(self, q, probes, k, dim_block, tau=None, running=True, threads=0):
        q, probes = f32(q), i64(probes)
        nq, P = probes.shape
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        st = OrcStats()
        keep = f32(tau) if tau is not None else None
        tp = _fp(keep) if keep is not None else None
        lib().orc_search_pruned(C.byref(self._s), nq, _fp(q), _ip(probes), P, k, dim_block, tp,
                                int(running), _fp(D), _ip(I), threads, C.byref(st))
        return D, I, st

    def warmup_tau(self, q, probes, k, L, W):
        q, probes = f32(q), i64(probes)
        nq, P = probes.shape
        tau = np.empty(nq, np.float32)
        lib().orc_warmup_tau(C.byref(self._s), nq, _fp(q), _ip(probes), P, k, L, W, _fp(tau))
        return tau

    def ref_search_base1(self, q, nprobe, k, L=10, W=0, threads=0):
        """Harmony base with one worker (oracle/_ref): warm-up bound, then the
        four-check pruned scan (L2)."""
        q = f32(q)
        nq = q.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        ref().ref_search_base1(nq, _fp(q), self.d, self.nlist, _fp(self.centroids), _ip(self.list_off),
                               _ip(self.ids), _fp(self.codes), nprobe, k, L, W if W > 0 else k, _fp(D), _ip(I),
                               threads)
        return D, I

    def ref_search_faiss(self, q, nprobe, k, threads=0):
        """The faiss IndexIVFFlat search algorithm (oracle/_ref, L2)."""
        q = f32(q)
        nq = q.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        ref().ref_search_faiss(nq, _fp(q), self.d, self.nlist, _fp(self.centroids), _ip(self.list_off),
                               _ip(self.ids), _fp(self.codes), nprobe, k, _fp(D), _ip(I), threads)
        return D, I

    def ref_search(self, q, nprobe, k, threads=0):
        """The reference's plain search (oracle/_ref)."""
        q = f32(q)
        nq = q.shape[0]
        P = self.nlist if (nprobe <= 0 or nprobe > self.nlist) else nprobe
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        probes = np.empty((nq, P), np.int64)
        ref().ref_search(self.metric, nq, _fp(q), self.d, self.nlist, _fp(self.centroids),
                         _ip(self.list_off), _ip(self.ids), _fp(self.codes), nprobe, k, _fp(D),
                         _ip(I), _ip(probes), threads)
        return D, I, probes
