"""CPU: the tuning table's host interface (magma_amd_tune_get / _set / _save /
_load).  Nothing here touches the device.  The timing side of the autotuner
(magma_amd_autotune) and every variant it can select are exercised on the
device by tests/test_gpu_variants.py."""
import os

import pytest


@pytest.fixture
def L():
    from magma_amd import lib
    L = lib.load()
    yield L
    for fam in range(3):
        for c in range(10):
            L.magma_amd_tune_set(fam, c, -1)


def test_defaults_are_untuned(L):
    for fam in range(3):
        for c in range(10):
            assert L.magma_amd_tune_get(fam, c) == -1


def test_set_rejects_out_of_range(L):
    assert L.magma_amd_tune_set(3, 0, 0) == -1           # no such family
    assert L.magma_amd_tune_set(0, 10, 0) == -2          # no such size class
    assert L.magma_amd_tune_set(0, 5, 3) == -3           # LU above order 128 has variants 0..2
    assert L.magma_amd_tune_set(0, 4, 5) == -3           # up to order 128, 0..4
    assert L.magma_amd_tune_set(0, 4, 4) == 0
    assert L.magma_amd_tune_set(2, 5, 4) == -3           # GEMM has variants 0..3
    assert L.magma_amd_tune_get(0, 5) == -1
    assert L.magma_amd_tune_get(7, 0) == -1


def test_save_load_round_trip(L, tmp_path):
    entries = {(0, 4): 2, (0, 5): 1, (1, 5): 2, (1, 6): 1, (2, 9): 1, (2, 2): 0}
    for (f, c), v in entries.items():
        assert L.magma_amd_tune_set(f, c, v) == 0
    path = str(tmp_path / "table.txt")
    assert L.magma_amd_tune_save(path.encode()) == 0
    text = open(path).read().splitlines()
    assert text[0].startswith("#")
    assert sorted(tuple(map(int, t.split())) for t in text[1:]) == sorted((f, c, v) for (f, c), v in entries.items())
    for (f, c) in entries:
        L.magma_amd_tune_set(f, c, -1)
    assert L.magma_amd_tune_load(path.encode()) == 0
    for (f, c), v in entries.items():
        assert L.magma_amd_tune_get(f, c) == v


def test_load_reports_bad_lines(L, tmp_path):
    path = str(tmp_path / "bad.txt")
    with open(path, "w") as f:
        f.write("# comment\n0 5 1\n0 5 9\n9 9 9\nnot a line\n")
    assert L.magma_amd_tune_load(path.encode()) == -104          # MAGMA_ERR_ILLEGAL_VALUE: two bad entries
    assert L.magma_amd_tune_get(0, 5) == 1
    assert L.magma_amd_tune_load(os.path.join(str(tmp_path), "missing").encode()) != 0
