"""The C++ mirror of harmony::Index (include/hivf_index.hpp) compiles and
links against libhivf.so (CPU), and a C++ caller shaped like the reference's
src/tests/index_test.cpp gets exact results on the GPU."""
import os
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "index_test.cpp")
LIBDIR = os.path.join(ROOT, "harmonyclient_amd")


def build(tmp_path):
    exe = tmp_path / "index_test"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}",
                    "-I/opt/rocm/include", SRC, "-o", str(exe), f"-L{LIBDIR}", "-lhivf",
                    f"-Wl,-rpath,{LIBDIR}"], check=True)
    return exe


def test_cpp_mirror_compiles_and_links(tmp_path):
    assert build(tmp_path).exists()


@pytest.mark.gpu
def test_cpp_caller_exact_on_gpu(tmp_path):
    print("[case] C++ caller through hivf_index.hpp", flush=True)
    exe = build(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout, r.stderr)
    assert r.returncode == 0


SHUFFLE_SRC = r"""
// The sampling of src/Clustering.cpp:48-89, as the reference writes it.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>
int main(int argc, char** argv) {
    size_t n = std::strtoull(argv[1], 0, 10), nlist = std::strtoull(argv[2], 0, 10);
    int seed = std::atoi(argv[3]);
    size_t max_samples = nlist * 256;
    std::vector<size_t> rows(n);
    std::iota(rows.begin(), rows.end(), 0);
    if (n > max_samples) {
        std::vector<size_t> indices(n);
        std::iota(indices.begin(), indices.end(), 0);
        std::shuffle(indices.begin(), indices.end(), std::default_random_engine(seed));
        rows.assign(indices.begin(), indices.begin() + max_samples);
        n = max_samples;
    }
    std::vector<size_t> indices(n);
    std::iota(indices.begin(), indices.end(), 0);
    std::shuffle(indices.begin(), indices.end(), std::default_random_engine(seed));
    std::printf("%zu\n", rows.size());
    for (size_t r : rows) std::printf("%zu ", r);
    std::printf("\n");
    for (size_t i = 0; i < nlist; ++i) std::printf("%zu ", indices[i]);
    std::printf("\n");
}
"""


@pytest.mark.parametrize("n,nlist,seed", [(5000, 16, 6666), (3000, 100, 6666), (70000, 64, 1234), (257, 1, 6666)])
def test_kmeans_sample_follows_the_reference_recipe(tmp_path, n, nlist, seed):
    """hivf_kmeans_sample picks the rows the reference's Clustering picks:
    a std::shuffle under std::default_random_engine(seed) for the
    subsample, then another for the initial centroids (compiled here with
    g++ and libstdc++ as the reference's build would be)."""
    import ctypes as C

    import numpy as np

    from harmonyclient_amd import index as hx

    src = tmp_path / "shuffle.cpp"
    src.write_text(SHUFFLE_SRC)
    exe = tmp_path / "shuffle"
    subprocess.run(["g++", "-std=c++17", "-O2", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(n), str(nlist), str(seed)], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    m = int(out[0])
    rows = np.array(out[1].split(), np.int64)
    init = np.array(out[2].split(), np.int64)
    cap = nlist * 256
    got_rows = np.empty(min(n, cap), np.int64)
    got_init = np.empty(nlist, np.int64)
    got_m = C.c_int64()
    i64 = C.POINTER(C.c_int64)
    rc = hx.lib().hivf_kmeans_sample(n, nlist, 256, seed, got_rows.ctypes.data_as(i64), C.byref(got_m),
                                     got_init.ctypes.data_as(i64))
    assert rc == 0
    assert got_m.value == m == min(n, cap)
    assert np.array_equal(got_rows, rows) and np.array_equal(got_init, init)
    assert len(set(init.tolist())) == nlist
