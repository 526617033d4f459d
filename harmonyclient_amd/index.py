"""Python host side of the hivf C ABI, mirroring harmony::Index.

``Index`` keeps the reference's method names and argument meaning
(src/Index.h:15-93): ``Index(d, nlist, nprobe, metric)``, ``train``, ``add``,
``search(n, queries, k, distances, labels) -> Stats``, ``save_index`` /
``load_index``.  Every call goes through ``libhivf.so`` (ctypes); there is no
fallback: a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhivf.so")

METRIC_IP = 0  # src/common.h:20-21
METRIC_L2 = 1

ALGO_PLAIN, ALGO_PRUNED, ALGO_BATCHED = 0, 1, 2
TAU_NONE, TAU_WARMUP, TAU_LEARNED = 0, 1, 2
PREC_FP32, PREC_BF16 = 0, 1

STATUS = {0: "OK", -1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ESTATE", -5: "EUNSUPPORTED",
          -6: "EIO"}

# Every symbol include/hivf.h declares.
EXPORTS = [
    "hivf_search_params_default", "hivf_index_create", "hivf_index_destroy", "hivf_last_error",
    "hivf_index_set_nprobe", "hivf_index_train", "hivf_index_set_centroids",
    "hivf_index_get_centroids", "hivf_index_add", "hivf_index_set_lists", "hivf_index_ntotal",
    "hivf_index_list_size", "hivf_index_info", "hivf_index_get_list", "hivf_index_search", "hivf_index_search_device",
    "hivf_index_coarse", "hivf_index_warmup_tau", "hivf_index_scan", "hivf_index_save_harmony",
    "hivf_index_load_harmony", "hivf_tau_model_create", "hivf_tau_model_destroy",
    "hivf_tau_model_num_params", "hivf_tau_model_get_params", "hivf_tau_model_set_params",
    "hivf_tau_model_forward", "hivf_tau_model_grad", "hivf_tau_model_train",
    "hivf_index_set_tau_model", "hivf_index_tau_traces", "hivf_index_tau_bounds", "hivf_index_autotune",
    "hivf_index_set_tune", "hivf_merge_topk", "hivf_merge_topk_device", "hivf_index_train_kmeans",
    "hivf_kmeans_sample", "hivf_serve_params_default", "hivf_server_start", "hivf_server_start_fn",
    "hivf_server_port", "hivf_server_wait", "hivf_server_stats_get", "hivf_server_error", "hivf_server_stop",
    "hivf_index_search_device_noalloc",
]


class HivfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class SearchParams(C.Structure):
    _fields_ = [
        ("nprobe", C.c_int64),
        ("nprobe_per_query", C.POINTER(C.c_int32)),
        ("algo", C.c_int32),
        ("tau_source", C.c_int32),
        ("precision", C.c_int32),
        ("dim_block", C.c_int32),
        ("warmup_lists", C.c_int32),
        ("warmup_list_size", C.c_int32),
        ("survivor_cap", C.c_int32),
        ("batch_queries", C.c_int32),
        ("batch_rows", C.c_int32),
        ("batch_waves", C.c_int32),
        ("use_graph", C.c_int32),
        ("reserved", C.c_int32),
        ("nprobe_per_query_dev", C.c_void_p),
    ]


class SearchStats(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double),
        ("coarse_ms", C.c_double),
        ("tau_ms", C.c_double),
        ("scan_ms", C.c_double),
        ("select_ms", C.c_double),
        ("fallback_ms", C.c_double),
        ("candidates", C.c_int64),
        ("blocks_total", C.c_int64),
        ("blocks_skipped", C.c_int64),
        ("survivors", C.c_int64),
        ("fallback_queries", C.c_int64),
        ("scan_launches", C.c_int64),
        ("exact_rescans", C.c_int64),
        ("batch_queries", C.c_int64),
        ("batch_rows", C.c_int64),
        ("batch_waves", C.c_int64),
        ("plan_ms", C.c_double),
        ("scan_kind", C.c_int64),
        ("overflow_queries", C.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class TuneResult(C.Structure):
    _fields_ = [
        ("batch_queries", C.c_int32),
        ("batch_rows", C.c_int32),
        ("batch_waves", C.c_int32),
        ("tried", C.c_int32),
        ("ms", C.c_double),
        ("ms_default", C.c_double),
        ("reserved", C.c_int32 * 4),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class ServeParams(C.Structure):
    _fields_ = [
        ("batch", C.c_int64),
        ("insert_phase", C.c_int32),
        ("train_niter", C.c_int32),
        ("max_request_rows", C.c_int64),
        ("search", SearchParams),
        ("reserved", C.c_int32 * 4),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("connections", C.c_int64),
        ("requests", C.c_int64),
        ("queries", C.c_int64),
        ("batches", C.c_int64),
        ("rows_inserted", C.c_int64),
        ("reserved", C.c_int64 * 3),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


SEARCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float),
                        C.POINTER(C.c_int64))
BUILD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_float))


class TrainParams(C.Structure):
    _fields_ = [
        ("epochs", C.c_int32),
        ("batch", C.c_int32),
        ("lr", C.c_float),
        ("quantile", C.c_float),
        ("miss_rate", C.c_float),
        ("holdout_every", C.c_int32),
        ("seed", C.c_uint64),
        ("target_scale", C.c_float),
        ("reserved", C.c_int32 * 3),
    ]


class TrainStats(C.Structure):
    _fields_ = [
        ("final_loss", C.c_double),
        ("holdout_miss_rate", C.c_double),
        ("holdout_mean_abs_err", C.c_double),
        ("offset", C.c_float),
        ("steps", C.c_int64),
        ("holdout", C.c_int64),
        ("reserved", C.c_int64 * 4),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


_f32p = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_lib = None


def lib():
    """Load libhivf.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run harmonyclient_amd.build.build()")
        # One HIP runtime per process.  torch ships its own libamdhip64 with
        # the same soname as /opt/rocm's; loading torch first makes
        # libhivf.so bind to that copy instead of starting a second runtime
        # that torch could then not initialise.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.hivf_search_params_default.argtypes = [C.POINTER(SearchParams)]
        L.hivf_index_create.restype = vp
        L.hivf_index_create.argtypes = [C.c_int64, C.c_int64, C.c_int]
        L.hivf_index_destroy.argtypes = [vp]
        L.hivf_last_error.restype = C.c_char_p
        L.hivf_last_error.argtypes = [vp]
        L.hivf_index_set_nprobe.argtypes = [vp, C.c_int64]
        L.hivf_index_train.argtypes = [vp, C.c_int64, _f32p, _f32p, C.c_int32]
        L.hivf_index_train_kmeans.argtypes = [vp, C.c_int64, _f32p, C.c_int32, C.c_uint64, C.c_int32]
        L.hivf_kmeans_sample.argtypes = [C.c_int64, C.c_int64, C.c_int32, C.c_uint64, _i64p, C.POINTER(C.c_int64),
                                         _i64p]
        L.hivf_index_set_centroids.argtypes = [vp, _f32p]
        L.hivf_index_get_centroids.argtypes = [vp, _f32p]
        L.hivf_index_add.argtypes = [vp, C.c_int64, _f32p]
        L.hivf_index_set_lists.argtypes = [vp, _i64p, _i64p, _f32p]
        L.hivf_index_ntotal.restype = C.c_int64
        L.hivf_index_ntotal.argtypes = [vp]
        L.hivf_index_list_size.argtypes = [vp, C.c_int64, _i64p]
        L.hivf_index_get_list.argtypes = [vp, C.c_int64, _i64p, _f32p]
        L.hivf_index_search.argtypes = [vp, C.c_int64, _f32p, C.c_int64, _f32p, _i64p,
                                        C.POINTER(SearchParams), C.POINTER(SearchStats)]
        L.hivf_index_search_device.argtypes = [vp, C.c_int64, vp, C.c_int64, vp, vp,
                                               C.POINTER(SearchParams), C.POINTER(SearchStats), vp]
        L.hivf_index_coarse.argtypes = [vp, C.c_int64, _f32p, C.c_int64, _f32p, _i64p]
        L.hivf_index_warmup_tau.argtypes = [vp, C.c_int64, _f32p, _i64p, C.c_int64, C.c_int64,
                                            C.c_int64, C.c_int64, _f32p]
        L.hivf_index_scan.argtypes = [vp, C.c_int64, _f32p, _i64p, C.c_int64, C.c_int64, _f32p,
                                      C.POINTER(SearchParams), _f32p, _i64p,
                                      C.POINTER(SearchStats)]
        L.hivf_index_save_harmony.argtypes = [vp, C.c_char_p]
        L.hivf_index_load_harmony.restype = vp
        L.hivf_index_load_harmony.argtypes = [C.c_char_p, _i32p]
        L.hivf_index_info.argtypes = [vp, _i64p, _i64p, C.POINTER(C.c_int)]
        L.hivf_tau_model_create.restype = vp
        L.hivf_tau_model_create.argtypes = [C.c_int32, C.c_int32, C.c_uint64]
        L.hivf_tau_model_destroy.argtypes = [vp]
        L.hivf_tau_model_num_params.restype = C.c_int64
        L.hivf_tau_model_num_params.argtypes = [vp]
        L.hivf_tau_model_get_params.argtypes = [vp, _f32p]
        L.hivf_tau_model_set_params.argtypes = [vp, _f32p]
        L.hivf_tau_model_forward.argtypes = [vp, C.c_int64, _f32p, _f32p]
        L.hivf_tau_model_grad.argtypes = [vp, C.c_int64, _f32p, _f32p, _f32p]
        L.hivf_tau_model_train.argtypes = [vp, C.c_int64, _f32p, _f32p, C.POINTER(TrainParams),
                                           C.POINTER(TrainStats)]
        L.hivf_index_set_tau_model.argtypes = [vp, vp]
        L.hivf_index_tau_traces.argtypes = [vp, C.c_int64, _f32p, C.c_int64, C.c_int64, _f32p, _f32p]
        L.hivf_index_tau_bounds.argtypes = [vp, C.c_int64, _f32p, C.c_int64, C.c_int64, _f32p]
        L.hivf_index_autotune.argtypes = [vp, C.c_int64, vp, C.c_int64, C.POINTER(SearchParams),
                                          C.c_int32, C.POINTER(TuneResult)]
        L.hivf_index_set_tune.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
        L.hivf_merge_topk.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int64, _f32p, _i64p,
                                      C.c_int64, _f32p, _i64p]
        L.hivf_merge_topk_device.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int64, vp, vp,
                                             C.c_int64, vp, vp, vp]
        L.hivf_index_search_device_noalloc.argtypes = [vp, C.c_int64, vp, C.c_int64, vp, vp,
                                                       C.POINTER(SearchParams), vp]
        L.hivf_serve_params_default.argtypes = [C.POINTER(ServeParams)]
        L.hivf_server_start.restype = vp
        L.hivf_server_start.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(ServeParams), C.POINTER(C.c_int)]
        L.hivf_server_start_fn.restype = vp
        L.hivf_server_start_fn.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_int64, C.c_int, SEARCH_FN, BUILD_FN,
                                           vp, C.POINTER(C.c_int)]
        L.hivf_server_port.argtypes = [vp]
        L.hivf_server_wait.argtypes = [vp]
        L.hivf_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
        L.hivf_server_error.restype = C.c_char_p
        L.hivf_server_error.argtypes = [vp]
        L.hivf_server_stop.argtypes = [vp]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(_f32p)


def _ip(a):
    return a.ctypes.data_as(_i64p)


def make_params(nprobe=0, algo=ALGO_PLAIN, tau=TAU_WARMUP, precision=PREC_FP32, dim_block=0,
                warmup_lists=10, warmup_list_size=0, nprobe_per_query=None, survivor_cap=0,
                batch_queries=0, batch_rows=0, batch_waves=0, use_graph=False, nprobe_per_query_dev=None):
    p = SearchParams()
    lib().hivf_search_params_default(C.byref(p))
    p.nprobe = nprobe
    p.algo = algo
    p.tau_source = tau
    p.precision = precision
    p.dim_block = dim_block
    p.warmup_lists = warmup_lists
    p.warmup_list_size = warmup_list_size
    p.survivor_cap = survivor_cap
    p.batch_queries = batch_queries
    p.batch_rows = batch_rows
    p.batch_waves = batch_waves
    p.use_graph = int(bool(use_graph))
    p.nprobe_per_query_dev = nprobe_per_query_dev
    if nprobe_per_query is not None:
        arr = np.ascontiguousarray(nprobe_per_query, dtype=np.int32)
        p._keep = arr
        p.nprobe_per_query = arr.ctypes.data_as(_i32p)
    return p


@dataclass
class Stats:
    """Counterpart of the reference's Stats (src/stats.h:14-152) for one search."""
    total_ms: float = 0.0
    coarse_ms: float = 0.0
    tau_ms: float = 0.0
    scan_ms: float = 0.0
    select_ms: float = 0.0
    fallback_ms: float = 0.0
    candidates: int = 0
    blocks_total: int = 0
    blocks_skipped: int = 0
    survivors: int = 0
    fallback_queries: int = 0
    scan_launches: int = 0
    exact_rescans: int = 0
    batch_queries: int = 0
    batch_rows: int = 0
    batch_waves: int = 0
    plan_ms: float = 0.0
    scan_kind: int = 0
    overflow_queries: int = 0

    @property
    def pruning_rate(self):
        return self.blocks_skipped / self.blocks_total if self.blocks_total else 0.0


class Index:
    """harmony::Index on the GPU (src/Index.h:15-93)."""

    def __init__(self, d, nlist, nprobe=1, metric=METRIC_L2, _handle=None):
        L = lib()
        self.d, self.nlist, self.metric = int(d), int(nlist), int(metric)
        h = _handle if _handle is not None else L.hivf_index_create(self.d, self.nlist, self.metric)
        if not h:
            raise HivfError(-1, L.hivf_last_error(None).decode())
        self._h = C.c_void_p(h)
        self.nprobe = nprobe

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hivf_index_destroy(h)
            self._h = None

    def _check(self, rc):
        if rc != 0:
            raise HivfError(rc, lib().hivf_last_error(self._h).decode())

    @property
    def nprobe(self):
        return self._nprobe

    @nprobe.setter
    def nprobe(self, v):
        self._nprobe = int(v)
        self._check(lib().hivf_index_set_nprobe(self._h, self._nprobe))

    # -- build -----------------------------------------------------------
    def train(self, n, codes, init_centroids=None, niter=10, seed=6666, max_points_per_centroid=256):
        """Lloyd k-means (src/Index.cpp:381-425).  Without init_centroids, the
        reference's recipe picks the sample and the initial centroids
        (src/Clustering.cpp:48-89, std::shuffle under
        std::default_random_engine(seed)); with them, every row of `codes` is
        a training row."""
        codes = _f(codes)[:n]
        if init_centroids is None:
            self._check(lib().hivf_index_train_kmeans(self._h, codes.shape[0], _fp(codes), niter, seed,
                                                      max_points_per_centroid))
            return
        init = _f(init_centroids)
        self._check(lib().hivf_index_train(self._h, codes.shape[0], _fp(codes), _fp(init), niter))

    def set_centroids(self, centroids):
        c = _f(centroids)
        assert c.shape == (self.nlist, self.d)
        self._check(lib().hivf_index_set_centroids(self._h, _fp(c)))

    @property
    def centroids(self):
        c = np.empty((self.nlist, self.d), np.float32)
        self._check(lib().hivf_index_get_centroids(self._h, _fp(c)))
        return c

    def add(self, n, codes):
        codes = _f(codes)[:n]
        self._check(lib().hivf_index_add(self._h, codes.shape[0], _fp(codes)))

    def set_lists(self, list_off, ids, codes):
        off = np.ascontiguousarray(list_off, np.int64)
        ids = np.ascontiguousarray(ids, np.int64)
        codes = _f(codes)
        self._check(lib().hivf_index_set_lists(self._h, _ip(off), _ip(ids), _fp(codes)))

    @property
    def ntotal(self):
        return lib().hivf_index_ntotal(self._h)

    def list_size(self, l):
        s = C.c_int64()
        self._check(lib().hivf_index_list_size(self._h, l, C.byref(s)))
        return s.value

    def get_list(self, l):
        n = self.list_size(l)
        ids = np.empty(n, np.int64)
        codes = np.empty((n, self.d), np.float32)
        self._check(lib().hivf_index_get_list(self._h, l, _ip(ids), _fp(codes)))
        return ids, codes

    def csr(self):
        """All lists as (list_off, ids, codes) in list order."""
        sizes = [self.list_size(l) for l in range(self.nlist)]
        off = np.zeros(self.nlist + 1, np.int64)
        off[1:] = np.cumsum(sizes)
        ids = np.empty(off[-1], np.int64)
        codes = np.empty((off[-1], self.d), np.float32)
        for l in range(self.nlist):
            if sizes[l]:
                ids[off[l]:off[l + 1]], codes[off[l]:off[l + 1]] = self.get_list(l)
        return off, ids, codes

    # -- search ----------------------------------------------------------
    def search(self, n, queries, k, distances=None, labels=None, params=None):
        """Index::search (src/Index.h:82).  Fills distances [n x k] and
        labels [n x k] (allocated when None); returns (distances, labels,
        Stats)."""
        q = _f(queries)[:n]
        n = q.shape[0]
        D = distances if distances is not None else np.empty((n, k), np.float32)
        I = labels if labels is not None else np.empty((n, k), np.int64)
        assert D.dtype == np.float32 and I.dtype == np.int64 and D.flags.c_contiguous
        st = SearchStats()
        p = params if params is not None else make_params()
        self._check(lib().hivf_index_search(self._h, n, _fp(q), k, _fp(D), _ip(I), C.byref(p),
                                            C.byref(st)))
        return D, I, Stats(**st.as_dict())

    def search_device(self, n, q_ptr, k, d_ptr, i_ptr, params=None, stats=True, stream=None):
        """Search on device buffers (raw pointers, e.g. torch data_ptr())."""
        st = SearchStats()
        p = params if params is not None else make_params()
        self._check(lib().hivf_index_search_device(self._h, n, q_ptr, k, d_ptr, i_ptr, C.byref(p),
                                                   C.byref(st) if stats else None, stream))
        return Stats(**st.as_dict()) if stats else None

    def coarse(self, queries, nprobe):
        q = _f(queries)
        P = self.nlist if (nprobe <= 0 or nprobe > self.nlist) else nprobe
        dist = np.empty((q.shape[0], P), np.float32)
        lists = np.empty((q.shape[0], P), np.int64)
        self._check(lib().hivf_index_coarse(self._h, q.shape[0], _fp(q), nprobe, _fp(dist),
                                            _ip(lists)))
        return dist, lists

    def warmup_tau(self, queries, probes, k, L=10, W=0):
        q = _f(queries)
        pr = np.ascontiguousarray(probes, np.int64)
        tau = np.empty(q.shape[0], np.float32)
        self._check(lib().hivf_index_warmup_tau(self._h, q.shape[0], _fp(q), _ip(pr), pr.shape[1],
                                                k, L, W if W > 0 else k, _fp(tau)))
        return tau

    def scan(self, queries, probes, k, algo=ALGO_PLAIN, tau=None, dim_block=0):
        q = _f(queries)
        pr = np.ascontiguousarray(probes, np.int64)
        n = q.shape[0]
        D = np.empty((n, k), np.float32)
        I = np.empty((n, k), np.int64)
        st = SearchStats()
        p = make_params(algo=algo, dim_block=dim_block)
        t = _f(tau) if tau is not None else None
        self._check(lib().hivf_index_scan(self._h, n, _fp(q), _ip(pr), pr.shape[1], k,
                                          _fp(t) if t is not None else None, C.byref(p), _fp(D),
                                          _ip(I), C.byref(st)))
        return D, I, Stats(**st.as_dict())

    # -- persistence -----------------------------------------------------
    def save_index(self, path):
        self._check(lib().hivf_index_save_harmony(self._h, str(path).encode()))

    @classmethod
    def load_index(cls, path, nprobe=1):
        """Index::load_index (src/Index.cpp:1735-1770)."""
        L = lib()
        st = C.c_int32()
        h = L.hivf_index_load_harmony(str(path).encode(), C.byref(st))
        if not h:
            raise HivfError(st.value, L.hivf_last_error(None).decode())
        d, nl, m = C.c_int64(), C.c_int64(), C.c_int()
        L.hivf_index_info(C.c_void_p(h), C.byref(d), C.byref(nl), C.byref(m))
        return cls(d.value, nl.value, nprobe, m.value, _handle=h)


N_FEATURES = 33  # 32 nearest coarse distances + query norm


class TauModel:
    """Learned pruning bound (MLP 33 -> hidden -> hidden -> 1) on the GPU."""

    def __init__(self, hidden=128, seed=1, in_dim=N_FEATURES):
        self.in_dim, self.hidden = in_dim, hidden
        h = lib().hivf_tau_model_create(in_dim, hidden, seed)
        if not h:
            raise HivfError(-3, "cannot create bound model")
        self._h = C.c_void_p(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hivf_tau_model_destroy(h)
            self._h = None

    @property
    def num_params(self):
        return lib().hivf_tau_model_num_params(self._h)

    def get_params(self):
        p = np.empty(self.num_params, np.float32)
        _chk(lib().hivf_tau_model_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        p = _f(p)
        assert p.size == self.num_params
        _chk(lib().hivf_tau_model_set_params(self._h, _fp(p)))

    def forward(self, x):
        x = _f(x)
        y = np.empty(x.shape[0], np.float32)
        _chk(lib().hivf_tau_model_forward(self._h, x.shape[0], _fp(x), _fp(y)))
        return y

    def grad(self, x, dy):
        x, dy = _f(x), _f(dy)
        g = np.empty(self.num_params, np.float32)
        _chk(lib().hivf_tau_model_grad(self._h, x.shape[0], _fp(x), _fp(dy), _fp(g)))
        return g

    def train(self, x, y, epochs=30, batch=1024, lr=2e-3, quantile=0.9, miss_rate=0.01,
              holdout_every=10, seed=7, target_scale=1.0):
        """Targets y = log1p(k-th distance / target_scale)."""
        x, y = _f(x), _f(y)
        tp = TrainParams(epochs, batch, lr, quantile, miss_rate, holdout_every, seed, target_scale)
        ts = TrainStats()
        _chk(lib().hivf_tau_model_train(self._h, x.shape[0], _fp(x), _fp(y), C.byref(tp),
                                        C.byref(ts)))
        return ts.as_dict()


def _chk(rc):
    if rc != 0:
        raise HivfError(rc, "bound model call failed")


def _index_tau_traces(self, queries, k, nprobe, chunk=50000):
    """Exact search of `queries`; returns (features [n x 33], k-th distance [n]).
    Searched `chunk` queries at a time (the batched search's scratch grows
    with the batch)."""
    q = _f(queries)
    F = np.empty((q.shape[0], N_FEATURES), np.float32)
    kth = np.empty(q.shape[0], np.float32)
    for s in range(0, q.shape[0], chunk):
        e = min(q.shape[0], s + chunk)
        qs = np.ascontiguousarray(q[s:e])
        Fs = np.empty((e - s, N_FEATURES), np.float32)
        ks = np.empty(e - s, np.float32)
        self._check(lib().hivf_index_tau_traces(self._h, e - s, _fp(qs), k, nprobe, _fp(Fs), _fp(ks)))
        F[s:e], kth[s:e] = Fs, ks
    return F, kth


def _index_tau_bounds(self, queries, k, nprobe):
    """The learned bound the batched search would use for `queries` [n]."""
    q = _f(queries)
    tau = np.empty(q.shape[0], np.float32)
    self._check(lib().hivf_index_tau_bounds(self._h, q.shape[0], _fp(q), k, nprobe, _fp(tau)))
    return tau


def _index_set_tau_model(self, model):
    self._check(lib().hivf_index_set_tau_model(self._h, model._h if model is not None else None))


def _index_train_tau_model(self, queries, k, nprobe, hidden=128, seed=1, **kw):
    """Traces from the exact search of training queries, then GPU training
    of the bound model; attaches it and returns (model, train stats)."""
    F, kth = self.tau_traces(queries, k, nprobe)
    keep = kth < np.finfo(np.float32).max
    m = TauModel(hidden=hidden, seed=seed)
    # distance scale: a hundredth of the median k-th distance, so that the
    # target log1p(D_k / s) is close to log D_k over the workload's range
    med = float(np.median(kth[keep])) if keep.any() else 0.0
    scale = np.float32(med / 100.0) if med > 0 else np.float32(1.0)
    y = np.log1p(kth[keep].astype(np.float64) / float(scale)).astype(np.float32)
    st = m.train(F[keep], y, target_scale=float(scale), **kw)
    st["target_scale"] = float(scale)
    self.set_tau_model(m)
    return m, st


def _index_autotune(self, n, q_ptr, k, params=None, reps=3):
    """Time the batched scan's tile shapes on n device queries; keeps the
    fastest as this index's default and returns it."""
    r = TuneResult()
    p = params if params is not None else make_params(algo=ALGO_BATCHED)
    self._check(lib().hivf_index_autotune(self._h, n, q_ptr, k, C.byref(p), reps, C.byref(r)))
    return r.as_dict()


def _index_set_tune(self, batch_queries=0, batch_rows=0, batch_waves=0):
    self._check(lib().hivf_index_set_tune(self._h, batch_queries, batch_rows, batch_waves))


def merge_topk(metric, D_parts, I_parts, k):
    """Merge per-partition top-k rows [parts x n x k_in] into [n x k]
    (src/Index.cpp:1327-1343) on the GPU."""
    Dp = np.ascontiguousarray(D_parts, np.float32)
    Ip = np.ascontiguousarray(I_parts, np.int64)
    parts, n, k_in = Dp.shape
    assert Ip.shape == Dp.shape
    D = np.empty((n, k), np.float32)
    I = np.empty((n, k), np.int64)
    rc = lib().hivf_merge_topk(metric, n, parts, k_in, _fp(Dp), _ip(Ip), k, _fp(D), _ip(I))
    if rc != 0:
        raise HivfError(rc, lib().hivf_last_error(None).decode())
    return D, I


def merge_topk_device(metric, n, parts, k_in, dp_ptr, ip_ptr, k, d_ptr, i_ptr, stream=None):
    """Device-buffer merge (raw pointers, e.g. torch data_ptr())."""
    rc = lib().hivf_merge_topk_device(metric, n, parts, k_in, dp_ptr, ip_ptr, k, d_ptr, i_ptr, stream)
    if rc != 0:
        raise HivfError(rc, lib().hivf_last_error(None).decode())


Index.autotune = _index_autotune
Index.set_tune = _index_set_tune
Index.tau_traces = _index_tau_traces
Index.tau_bounds = _index_tau_bounds
Index.set_tau_model = _index_set_tau_model
Index.train_tau_model = _index_train_tau_model


def make_serve_params(batch=4096, insert_phase=False, train_niter=20, max_request_rows=0, search=None):
    p = ServeParams()
    lib().hivf_serve_params_default(C.byref(p))
    p.batch = batch
    p.insert_phase = int(bool(insert_phase))
    p.train_niter = train_niter
    p.max_request_rows = max_request_rows
    if search is not None:
        p.search = search
    return p


class Server:
    """The TCP serve front-end (include/hivf.h, hivf_server_*): the
    reference's INSERT / QUERY protocol in front of an index's captured
    search, or in front of a Python search function (``search_fn(x, k) ->
    (D, I)``, ``build_fn(x)``) for callers that bring their own."""

    def __init__(self, index=None, address="127.0.0.1", port=0, params=None, d=None, search_fn=None,
                 build_fn=None, batch=4096, insert_phase=False):
        st = C.c_int(0)
        L = lib()
        self._keep = []
        if index is not None:
            p = params if params is not None else make_serve_params()
            h = L.hivf_server_start(index._h, address.encode(), port, C.byref(p), C.byref(st))
            self._index = index
        else:
            dd = int(d)

            def _search(user, n, x, k, D, I):
                try:
                    xs = np.ctypeslib.as_array(x, shape=(n, dd))
                    Dv, Iv = search_fn(xs, k)
                    np.ctypeslib.as_array(D, shape=(n, k))[:] = Dv
                    np.ctypeslib.as_array(I, shape=(n, k))[:] = Iv
                    return 0
                except Exception:  # noqa: BLE001 - reported as a status code
                    return -1

            def _build(user, n, x):
                try:
                    build_fn(np.ctypeslib.as_array(x, shape=(n, dd)).copy() if n else np.zeros((0, dd), np.float32))
                    return 0
                except Exception:  # noqa: BLE001
                    return -1

            sf, bf = SEARCH_FN(_search), BUILD_FN(_build) if build_fn is not None else BUILD_FN(0)
            self._keep += [sf, bf]
            h = L.hivf_server_start_fn(dd, address.encode(), port, batch, int(bool(insert_phase)), sf, bf, None,
                                       C.byref(st))
        if not h:
            raise HivfError(st.value, L.hivf_last_error(None).decode())
        self._h = C.c_void_p(h)

    @property
    def port(self):
        return lib().hivf_server_port(self._h)

    def wait(self):
        return lib().hivf_server_wait(self._h)

    def stats(self):
        s = ServerStats()
        _chk(lib().hivf_server_stats_get(self._h, C.byref(s)))
        return s.as_dict()

    def error(self):
        return lib().hivf_server_error(self._h).decode()

    def stop(self):
        if self._h is not None and self._h.value:
            lib().hivf_server_stop(self._h)
            self._h = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hivf_server_stop(h)
