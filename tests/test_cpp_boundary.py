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
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
