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


def test_host_quad9_meshes_match_oracle():
    """Host quad9 mesh generators equal the oracle's (regular_mesh.py:107-145, :182-218)."""
    import numpy as np
    from jax_fem_cfd.setup import regular_mesh as rm
    from oracle import mesh as ms
    for f_host, f_or in ((rm.create_mesh_2d_quad9, ms.create_mesh_2d_quad9),
                         (rm.create_periodic_mesh_2d_quad9, ms.create_periodic_mesh_2d_quad9)):
        a, b = f_host(5, 3, 2.0, 1.5), f_or(5, 3, 2.0, 1.5)
        assert np.allclose(np.asarray(a[0]), b[0], rtol=0, atol=1e-15)
        assert np.array_equal(np.asarray(a[1]), b[1])
        if b[2] is not None:
            for x, y in zip(a[2], b[2]):
                assert np.array_equal(np.asarray(x), y)
            for x, y in zip(a[3], b[3]):
                assert np.array_equal(np.asarray(x), y)


def test_host_hex27_mesh_matches_oracle_bit_exact():
    """The host hex27 mesh generator (written separately) equals the oracle's bit for bit:
    coordinates, connectivity, face node lists, quad9 face patches and normals."""
    from jax_fem_cfd.setup import regular_mesh as rm
    from oracle import mesh as oms
    for n in [(3, 2, 2), (1, 1, 1), (4, 3, 2)]:
        a = oms.create_mesh_3d_hex27(*n, 1.5, 1.0, 0.5)
        b = rm.create_mesh_3d_hex27(*n, 1.5, 1.0, 0.5)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        for f in range(6):
            assert np.array_equal(a[2][f], b[2][f])
            assert np.array_equal(a[3][f], b[3][f])
            assert np.array_equal(a[4][f], b[4][f])

