"""CPU: the tuning cache file (ACLFFT_TUNE_CACHE) - plan-level choices of uneven plans written and
read back per group list and requested precision, including a cache path that is a bare file name
(created in the working directory)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ops-fft_amd")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    lib = os.path.join(PKG, "libcann_ops_fft.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    exe = str(tmp_path_factory.mktemp("tc") / "tune_cache_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "lib"),
                           os.path.join(ROOT, "tests", "cpp", "tune_cache_check.cpp"),
                           os.path.join(PKG, "lib", "tune_cache.cpp"), lib, "-Wl,-rpath," + PKG, "-o", exe])
    return exe


@pytest.mark.parametrize("bare", [True, False])
def test_plan_choice_round_trip(checker, tmp_path, bare):
    path = "tune.txt" if bare else str(tmp_path / "sub" / "dir" / "tune.txt")
    env = dict(os.environ, ACLFFT_TUNE_CACHE=path)
    out = subprocess.run([checker], cwd=str(tmp_path), env=env, capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    f = tmp_path / "tune.txt" if bare else tmp_path / "sub" / "dir" / "tune.txt"
    assert f.is_file()
    text = f.read_text()
    assert " plan auto 1:0:3:1:2" in text
