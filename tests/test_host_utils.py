"""Solution / iteration-history files (jax_fem_cfd/utils/running.py:16-38): the
.npy layouts are [len(U), U..., p...] and a (3, n) stack (momentum, pressure,
steps) -- checked against arrays built here by hand, and round-tripped."""
import numpy as np

from jax_fem_cfd.utils import running


def test_solution_file_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    U, p = rng.standard_normal(10), rng.standard_normal(5)
    path = tmp_path / "sol.npy"
    running.save_sol(U, p, str(path))
    raw = np.load(path)
    assert np.array_equal(raw, np.concatenate([[10.0], U, p]))
    U2, p2 = running.load_sol(str(path))
    assert np.array_equal(U2, U) and np.array_equal(p2, p)


def test_iteration_history_layout_and_round_trip(tmp_path):
    m, pr, st = np.array([3, 4, 5]), np.array([10, 9, 8]), np.array([1, 2, 3])
    path = tmp_path / "hist.npy"
    running.save_iter_hist(m, pr, st, str(path))
    assert np.array_equal(np.load(path), np.stack([m, pr, st]))
    a, b, c = running.load_iter_hist(str(path))
    assert np.array_equal(a, m) and np.array_equal(b, pr) and np.array_equal(c, st)
