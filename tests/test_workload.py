"""Workload generator: Philox4x32-10 known-answer vectors (Random123's
published kat_vectors) and distribution sanity of the synthetic data."""
import numpy as np

from culsh_amd import workload as wl


def test_philox_known_answers():
    assert wl.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C,
                                                      0x9B00DBD8]
    assert wl.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E,
                                                                  0xA20BC7C6, 0x6D5451FD]
    assert wl.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                            [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                          0x24126EA1]


def test_gamma_ints_distribution():
    X = wl.gamma_ints(20000, 32, 5)
    assert X.dtype == np.float32 and X.shape == (20000, 32)
    assert np.all(X == np.round(X)) and X.min() == 0 and X.max() <= 255
    assert abs((X == 0).mean() - 0.6) < 0.01
    nz = X[X > 0]
    assert abs(nz.mean() - 50.0) < 1.5          # Gamma(2, 25) mean


def test_mixture_and_queries():
    X = wl.mixture(5000, 16, 10, 0.1, 3)
    assert np.all(np.abs(X) < 1.8)
    Y = wl.queries_near(X, 100, 0.0, 4)
    # noise 0: every query is a dataset vector
    assert all((X == y).all(axis=1).any() for y in Y[:20])


def test_projection_parameters():
    A, b = wl.projection(4, 64, 8, 2.0, 9)
    assert abs(A.mean()) < 0.05 and abs(A.std() - 1.0) < 0.05
    assert b.min() >= 0 and b.max() < 2.0


def test_generators_are_deterministic():
    assert np.array_equal(wl.gamma_ints(100, 8, 1), wl.gamma_ints(100, 8, 1))
    assert not np.array_equal(wl.gamma_ints(100, 8, 1), wl.gamma_ints(100, 8, 2))
