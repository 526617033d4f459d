"""Index and result file formats on the CPU.

* Malformed index files (truncated anywhere, list sizes beyond the file,
  sub_k beyond the file, ids of another dimension, list ids beyond nlist)
  make hivf_index_load_harmony return NULL with HIVF_EIO -- never an
  exception or an abort across the C ABI.  A well-formed file written by the
  independent writer in tests/reffile.py passes the parser; on a machine
  without a GPU the load then stops at index creation (HIVF_EHIP).
* The results ``.bin`` format (src/utils.h:565-620) round-trips through
  harmonyclient_amd.fvecs and matches the byte layout the reference writes.
* A host build of the parser and the oracle under AddressSanitizer and
  UndefinedBehaviorSanitizer runs clean (tests/cpp/sanitize_main.cpp).
"""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

import reffile
from conftest import ROOT, gpu_available
from harmonyclient_amd import fvecs
from harmonyclient_amd import index as hx

EIO, EHIP = -6, -3


def _load(path):
    st = C.c_int32(123)
    h = hx.lib().hivf_index_load_harmony(str(path).encode(), C.byref(st))
    return h, st.value


def _small_file(path, opt_of=None):
    rng = np.random.default_rng(0)
    d, nlist = 8, 5
    cent = rng.integers(0, 255, (nlist, d)).astype(np.float32)
    lists = {}
    row = 0
    for l in (0, 1, 3, 4):  # list 2 empty
        n = 3 + l
        lists[l] = (np.arange(row, row + n), rng.integers(0, 255, (n, d)).astype(np.float32))
        row += n
    reffile.write_index(path, d, nlist, 1, cent, lists, opt_of=opt_of, sub_k=2)
    return d, nlist, cent, lists


def _expect_parsed(st):
    # the parser accepted the file: without a GPU the load stops at creating
    # the index
    if gpu_available():
        assert st == 0
    else:
        assert st == EHIP
        assert b"cannot create index" in hx.lib().hivf_last_error(None)


def test_wellformed_file_passes_the_parser(tmp_path):
    p = tmp_path / "ok.index"
    _small_file(p, {1: reffile.OPT_TRIANGLE, 3: reffile.OPT_SUBNN_L2, 4: reffile.OPT_SUBNN_IP | reffile.OPT_TRIANGLE})
    h, st = _load(p)
    _expect_parsed(st)
    if h:
        hx.lib().hivf_index_destroy(h)


def test_every_truncation_is_eio_or_a_shorter_index(tmp_path):
    p = tmp_path / "full.index"
    _small_file(p, {1: reffile.OPT_TRIANGLE, 3: reffile.OPT_SUBNN_L2})
    raw = p.read_bytes()
    header = 48 + 5 * 8 * 4
    # record boundaries: a cut there (or inside a list id) is a shorter valid
    # file, as the reference's reader stops at the first list id it cannot read
    ok_cuts = set()
    pos = header
    while pos < len(raw):
        ok_cuts.update(range(pos, pos + 8))
        n, d, sk, opt = struct.unpack_from("<QQQi", raw, pos + 8)
        rec = 8 + 28 + n * (8 + 4 * d + 4)
        if opt & (1 | 4):
            rec += 8 * n
        if opt & 2:
            rec += n * sk * 12
        if opt & 4:
            rec += 2 * n * sk * 12
        pos += rec
    assert pos == len(raw)
    ok_cuts.add(len(raw))
    q = tmp_path / "cut.index"
    for cut in range(len(raw)):
        q.write_bytes(raw[:cut])
        h, st = _load(q)
        if cut in ok_cuts:
            _expect_parsed(st)
        else:
            assert not h and st == EIO, (cut, st)
        if h:
            hx.lib().hivf_index_destroy(h)


@pytest.mark.parametrize("field,value", [("n", 1 << 40), ("n", (1 << 64) - 1), ("n", 1 << 61), ("d", 9),
                                         ("sk", 1 << 62), ("lid", 5), ("lid", 1 << 63), ("opt", 8)])
def test_oversized_or_inconsistent_record_is_eio(tmp_path, field, value):
    d, nlist = 8, 5
    cent = np.zeros((nlist, d), np.float32)
    ids = np.arange(4)
    codes = np.ones((4, d), np.float32)
    rec = bytearray(reffile.list_record(1, ids, codes, opt=reffile.OPT_SUBNN_L2, sub_k=2))
    off = {"lid": 0, "n": 8, "d": 16, "sk": 24}
    if field == "opt":
        struct.pack_into("<i", rec, 32, value)
    else:
        struct.pack_into("<Q", rec, off[field], value)
    p = tmp_path / "bad.index"
    p.write_bytes(reffile.header(d, nlist, 1, cent) + bytes(rec))
    h, st = _load(p)
    assert not h and st == EIO, st


@pytest.mark.parametrize("hdr", [(0, 5, 1), (8, 0, 1), (8, 5, 3), (1 << 40, 5, 1), (8, 1 << 40, 1)])
def test_implausible_header_is_eio(tmp_path, hdr):
    d, nlist, metric = hdr
    p = tmp_path / "h.index"
    p.write_bytes(struct.pack("<QQiiQQQ", d, nlist, metric, 0, 0, 1, 1) + b"\0" * 64)
    h, st = _load(p)
    assert not h and st == EIO


def test_header_claiming_more_centroids_than_the_file_holds(tmp_path):
    p = tmp_path / "c.index"
    p.write_bytes(struct.pack("<QQiiQQQ", 128, 1 << 20, 1, 0, 0, 1, 1) + b"\0" * 1000)
    h, st = _load(p)
    assert not h and st == EIO
    assert b"truncated centroids" in hx.lib().hivf_last_error(None)


def test_results_bin_round_trip(tmp_path):
    """src/utils.h:585-593 writes int64 labels then float32 distances, no
    header; loadResults (src/utils.h:610-618) reads them back the same way."""
    rng = np.random.default_rng(1)
    nq, k = 7, 5
    I = rng.integers(-1, 1 << 40, (nq, k)).astype(np.int64)
    D = rng.random((nq, k)).astype(np.float32)
    p = tmp_path / "res.bin"
    fvecs.write_results(str(p), I, D)
    raw = p.read_bytes()
    assert len(raw) == nq * k * 12
    assert raw[: nq * k * 8] == I.tobytes() and raw[nq * k * 8:] == D.tobytes()
    I2, D2 = fvecs.read_results(str(p), nq, k)
    assert np.array_equal(I2, I) and np.array_equal(D2.view(np.uint32), D.view(np.uint32))
    with pytest.raises(ValueError):
        fvecs.read_results(str(p), nq + 1, k)


def test_sanitizer_build_runs_clean(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "sanitize_main.cpp")
    fmt = os.path.join(ROOT, "harmonyclient_amd", "csrc", "harmony_format.cpp")
    orc = os.path.join(ROOT, "oracle", "oracle.c")
    inc = [f"-I{os.path.join(ROOT, d)}" for d in ("include", "oracle", "harmonyclient_amd/csrc")]
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
    obj = tmp_path / "oracle.o"
    subprocess.run(["gcc", "-std=c11", "-ffp-contract=off", "-fopenmp", *san, *inc, "-c", orc, "-o", str(obj)],
                   check=True)
    exe = tmp_path / "sanitize"
    subprocess.run(["g++", "-std=c++17", "-ffp-contract=off", "-fopenmp", *san, *inc, src, fmt, str(obj),
                    "-o", str(exe), "-lm"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0
    assert "sanitize ok" in r.stdout
