"""CPU: the request server's memory bookkeeping (magma_amd/csrc/serve_cache.h,
the power-of-two arena and the resident-factorization cache that serve.hip
drives) built with AddressSanitizer and UBSan and run through
tests/serve_cache/stress.cpp: two regression scenarios (a repeat request
whose own resident fills the budget, which used to be refused for ever; two
block sizes that fit the free space alone but not together, which used to be
refused without evicting) and random admission / completion / eviction /
re-carving sequences checked after every step for overlapping, escaped or
double-handed blocks, byte accounting and pin counts."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "serve_cache", "stress.cpp")


def test_serve_cache_stress_asan_ubsan(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "stress")
    subprocess.run([gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "magma_amd", "csrc"), SRC, "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "self-pinned hit: ok" in r.stdout and "two sizes: ok" in r.stdout
    assert r.stdout.strip().endswith("ok")
