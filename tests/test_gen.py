"""The synthetic-input generator: determinism, distributions, and agreement of
the numpy and torch implementations (CPU)."""
import numpy as np
import pytest

from magma_amd import gen


def test_splitmix_reference_values():
    # splitmix64 of 0 (the first output of the generator seeded with 0)
    assert int(gen.splitmix64(np.uint64(0))) == 0xE220A8397B1DCDAF
    assert int(gen.splitmix64(np.uint64(1))) == 0x910A2DEC89025CC1


def test_uniform_range_and_determinism():
    a = gen.uniform_pm1(0, np.arange(4), 1000)
    b = gen.uniform_pm1(0, np.arange(4), 1000)
    assert np.array_equal(a, b)
    assert a.min() >= -1.0 and a.max() < 1.0
    assert abs(a.mean()) < 0.05
    assert not np.array_equal(gen.uniform_pm1(1, [0], 10), gen.uniform_pm1(0, [0], 10))


def test_normal_moments():
    z = gen.normal(1, np.arange(8), 20000).ravel()
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


def test_torch_matches_numpy():
    torch = pytest.importorskip("torch")
    a = gen.uniform_pm1(3, np.arange(5, 9), 777)
    t = gen.uniform_pm1_torch(3, 5, 4, 777, "cpu").numpy()
    assert np.array_equal(a, t)
    zn = gen.normal(3, np.arange(2), 500)
    zt = gen.normal_torch(3, 0, 2, 500, "cpu").numpy()
    assert np.allclose(zn, zt, rtol=1e-14, atol=1e-14)


def test_spd_and_cond():
    A = gen.spd_batch(1, 3, 16)
    for M in A:
        assert np.allclose(M, M.T)
        assert np.linalg.eigvalsh(M).min() > 0
    A, kappa = gen.cond_batch(3, 4, 32)
    for M, k in zip(A, kappa):
        c = np.linalg.cond(M.T)
        assert abs(np.log10(c) - np.log10(k)) < 1e-6
        assert 1e2 <= k <= 1e6


def test_vbatch_sizes():
    n = gen.vbatch_sizes(2, 50000, 8, 256)
    assert n.min() == 8 and n.max() == 256
    # log-uniform: about half the sizes below the geometric mean
    assert abs(np.mean(n < np.sqrt(8 * 257)) - 0.5) < 0.02
