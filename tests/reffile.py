"""Writer of the reference's index file format, independent of the library's
own writer (harmonyclient_amd/csrc/harmony_format.cpp), for load tests.

Layout, src/Index.cpp:1708-1732 (Index::save_index) and src/IVF.cpp:76-99
(IVF::save_IVF): size_t d, nlist; MetricType (int32); OptLevel (int32);
size_t sub_k, sub_nlist, sub_nprobe; float centroids[nlist*d]; then per
non-empty list: size_t list_id, size_t list_size, d, sub_k; OptLevel;
size_t ids[n]; float codes[n*d]; float norms[n]; then the opt level's
extras: candidate2centroid and its square root (TRIANGLE or SUBNN_IP),
sub-nearest L2 ids and distances (SUBNN_L2), sub-nearest and sub-farthest IP
ids and distances (SUBNN_IP).
"""
import struct

import numpy as np

OPT_NONE, OPT_TRIANGLE, OPT_SUBNN_L2, OPT_SUBNN_IP = 0, 1, 2, 4


def _norms_fp32(codes):
    """Sequential FP32 sum of squares per row (src/Index.cpp:689-696)."""
    out = np.zeros(codes.shape[0], np.float32)
    for c in range(codes.shape[1]):
        out = (out + codes[:, c] * codes[:, c]).astype(np.float32)
    return out


def list_record(list_id, ids, codes, opt=OPT_NONE, sub_k=0, seed=0, size_override=None):
    """Bytes of one (list id, IVF record) pair.  Extras are filled with
    seeded values: the loader must skip them."""
    ids = np.asarray(ids, np.uint64)
    codes = np.ascontiguousarray(codes, np.float32)
    n, d = codes.shape
    rng = np.random.default_rng(seed)
    out = [struct.pack("<QQQQi", list_id, n if size_override is None else size_override, d, sub_k, opt),
           ids.tobytes(), codes.tobytes(), _norms_fp32(codes).tobytes()]
    if opt & (OPT_TRIANGLE | OPT_SUBNN_IP):
        c2 = rng.random(n).astype(np.float32)
        out += [c2.tobytes(), np.sqrt(c2).astype(np.float32).tobytes()]
    if opt & OPT_SUBNN_L2:
        out += [rng.integers(0, 1 << 40, n * sub_k, dtype=np.int64).tobytes(),
                rng.random(n * sub_k).astype(np.float32).tobytes()]
    if opt & OPT_SUBNN_IP:
        for _ in range(2):
            out += [rng.integers(0, 1 << 40, n * sub_k, dtype=np.int64).tobytes(),
                    rng.random(n * sub_k).astype(np.float32).tobytes()]
    return b"".join(out)


def header(d, nlist, metric, centroids, opt=OPT_NONE, sub_k=0, sub_nlist=1, sub_nprobe=1):
    c = np.ascontiguousarray(centroids, np.float32)
    assert c.shape == (nlist, d)
    return struct.pack("<QQiiQQQ", d, nlist, metric, opt, sub_k, sub_nlist, sub_nprobe) + c.tobytes()


def write_index(path, d, nlist, metric, centroids, lists, opt_of=None, sub_k=3):
    """lists: {list_id: (ids, codes)} (empty lists are left out, as the
    reference does); opt_of: {list_id: opt level} for the records."""
    opt_of = opt_of or {}
    with open(path, "wb") as f:
        f.write(header(d, nlist, metric, centroids))
        for lid in sorted(lists):
            ids, codes = lists[lid]
            if len(ids) == 0:
                continue
            f.write(list_record(lid, ids, codes, opt_of.get(lid, OPT_NONE), sub_k, seed=lid))
