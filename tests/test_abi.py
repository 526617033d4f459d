"""C ABI checks that need no GPU: libhivf.so loads, exports every function
include/hivf.h declares, its structs have the layout the Python host
assumes, and calls fail with status codes (never crash) where no device is
usable."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, gpu_available
from harmonyclient_amd import index as hx

HEADER = os.path.join(ROOT, "include", "hivf.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hivf_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = hx.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(hx.EXPORTS) == names


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "hivf.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\","
        "sizeof(hivf_search_params), offsetof(hivf_search_params, algo),"
        "offsetof(hivf_search_params, warmup_list_size), sizeof(hivf_search_stats),"
        "offsetof(hivf_search_stats, candidates), offsetof(hivf_search_stats, plan_ms),"
        "offsetof(hivf_search_params, use_graph), sizeof(hivf_tune_result),"
        "offsetof(hivf_tune_result, ms_default));"
        "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)],
                   check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True)
                   .stdout.split()))
    want = [C.sizeof(hx.SearchParams), hx.SearchParams.algo.offset,
            hx.SearchParams.warmup_list_size.offset, C.sizeof(hx.SearchStats),
            hx.SearchStats.candidates.offset, hx.SearchStats.plan_ms.offset,
            hx.SearchParams.use_graph.offset, C.sizeof(hx.TuneResult), hx.TuneResult.ms_default.offset]
    assert got == want


def test_header_compiles_as_c_and_cxx(tmp_path):
    src = tmp_path / "use.c"
    src.write_text('#include "hivf.h"\nint f(hivf_index* h){return hivf_index_set_nprobe(h, 4);}\n')
    for cc in (["gcc", "-std=c99"], ["g++", "-x", "c++", "-std=c++17"]):
        subprocess.run([*cc, "-Wall", "-Werror", "-c", f"-I{os.path.join(ROOT, 'include')}",
                        str(src), "-o", str(tmp_path / "use.o")], check=True)


def test_defaults():
    p = hx.make_params()
    assert p.algo == hx.ALGO_PLAIN and p.warmup_lists == 10 and p.nprobe == 0


def test_null_handle_calls_return_einval():
    lib = hx.lib()
    assert lib.hivf_index_set_nprobe(None, 4) == -1
    assert lib.hivf_index_search(None, 1, None, 1, None, None, None, None) == -1
    assert lib.hivf_index_ntotal(None) == 0
    lib.hivf_index_destroy(None)


def test_bad_shape_rejected_before_device():
    lib = hx.lib()
    assert not lib.hivf_index_create(0, 10, 1)
    assert b"bad d" in lib.hivf_last_error(None)
    assert not lib.hivf_index_create(16, 10, 7)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU failure path")
def test_create_without_gpu_fails_cleanly():
    with pytest.raises(hx.HivfError):
        hx.Index(128, 16)


def test_load_missing_file_reports_eio(tmp_path):
    lib = hx.lib()
    st = C.c_int32()
    h = lib.hivf_index_load_harmony(str(tmp_path / "none.index").encode(), C.byref(st))
    assert not h and st.value == -6


def test_merge_argument_checks_without_device():
    lib = hx.lib()
    z = np.zeros(1, np.float32)
    zi = np.zeros(1, np.int64)
    fp = z.ctypes.data_as(C.POINTER(C.c_float))
    ip = zi.ctypes.data_as(C.POINTER(C.c_int64))
    # empty batch: nothing to do, no device needed
    assert lib.hivf_merge_topk(hx.METRIC_L2, 0, 2, 5, fp, ip, 3, fp, ip) == 0
    assert lib.hivf_merge_topk(7, 1, 1, 1, fp, ip, 1, fp, ip) == -1          # bad metric
    assert lib.hivf_merge_topk(hx.METRIC_L2, 1, 0, 1, fp, ip, 1, fp, ip) == -1  # no partitions
    assert lib.hivf_merge_topk(hx.METRIC_L2, 1, 1, 1, fp, ip, 2000, fp, ip) == -5  # k > 1024
    assert b"k <= 1024" in lib.hivf_last_error(None)
    assert lib.hivf_merge_topk(hx.METRIC_L2, 1, 1, 1, None, ip, 1, fp, ip) == -1
    zi[0] = 1 << 33
    assert lib.hivf_merge_topk(hx.METRIC_L2, 1, 1, 1, fp, ip, 1, fp, ip) == -5  # id beyond 32 bits
    assert lib.hivf_merge_topk_device(hx.METRIC_L2, 4, 1, 1, None, None, 1, None, None, None) == -1


def test_tuning_calls_reject_null_handle():
    lib = hx.lib()
    assert lib.hivf_index_set_tune(None, 32, 1024, 16) == -1
    r = hx.TuneResult()
    assert lib.hivf_index_autotune(None, 10, None, 10, None, 1, C.byref(r)) == -1
