"""CPU-side checks of the product: the C-ABI library loads and exports every
entry point include/fcfd.h declares; the host set-up logic (Config
validation, mesh index arrays, BC index sets, grid recognition) matches the
oracle bit for bit.  No GPU compute here."""
import os
import re

import numpy as np
import pytest

from oracle import mesh as oms
from oracle import ns

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    src = open(os.path.join(ROOT, "include", "fcfd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fcfd_\w+)\s*\(", src, re.M)))


def test_library_exports_every_header_symbol():
    from jax_fem_cfd import _native
    lib = _native.load_library()
    declared = _header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_native.exported_symbols()) == declared


def test_library_missing_fails_loudly(tmp_path):
    from jax_fem_cfd import _native
    with pytest.raises(_native.NativeError):
        _native.load_library(str(tmp_path / "libfcfd.so"))


@pytest.mark.parametrize("nex,ney,periodic", [(5, 3, False), (8, 8, False), (6, 4, True)])
def test_host_mesh_matches_oracle_bit_exact(nex, ney, periodic):
    from jax_fem_cfd.setup import regular_mesh as rm
    f = rm.create_periodic_mesh_2d if periodic else rm.create_mesh_2d
    g = oms.create_periodic_mesh_2d if periodic else oms.create_mesh_2d
    a = f(nex, ney, 1.3, 0.7)
    b = g(nex, ney, 1.3, 0.7)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    if not periodic:
        for k in range(4):
            assert np.array_equal(a[2][k], b[2][k])
            assert np.array_equal(a[3][k], b[3][k])
            assert np.array_equal(a[4][k], b[4][k])


def test_host_mesh_3d_matches_oracle():
    from jax_fem_cfd.setup import regular_mesh as rm
    a = rm.create_mesh_3d(3, 2, 2, 1.0, 1.0, 1.0)
    b = oms.create_mesh_3d(3, 2, 2, 1.0, 1.0, 1.0)
    assert np.array_equal(a[1], b[1]) and np.allclose(a[0], b[0])
    for k in range(6):
        assert np.array_equal(a[3][k], b[3][k])


def test_config_validation_errors():
    import jax_fem_cfd.setup.simulation as sim
    mc = sim.SimpleMeshConfig([16, 16], [1, 1], 1, False)
    sc = sim.StandardSolverConfig([0], [], [], False, 1e-8, True, "multigrid", None)
    with pytest.raises(ValueError, match="final_multigrid_mesh"):
        sim.Config("simple", "standard", mc, sc)
    sc = sim.StandardSolverConfig([0], [], [], False, 1e-8, True, "multigrid", [5, 5])
    with pytest.raises(ValueError, match="log2"):
        sim.Config("simple", "standard", mc, sc)
    sc = sim.StandardSolverConfig([0], [], [], False, 1e-8, True, "multigrid", [32, 32])
    with pytest.raises(ValueError, match="too large"):
        sim.Config("simple", "standard", mc, sc)
    with pytest.raises(ValueError, match="Unknown mesh"):
        sim.Config("other", "standard", mc, sim.StandardSolverConfig([0], [], [], False, 1e-8, True))
    with pytest.raises(ValueError, match="precision"):
        sim.Config("simple", "standard", mc, sim.StandardSolverConfig([0], [], [], False, 1e-8, True,
                                                                      precision="fp8"))
    sim.Config("simple", "standard", mc, sim.StandardSolverConfig([0], [], [], False, 1e-8, True, "multigrid",
                                                                  [4, 4]))


def test_simple_mesh_bc_sets_match_oracle():
    import jax_fem_cfd.setup.simulation as sim
    mc = sim.SimpleMeshConfig([8, 6], [1.0, 0.75], 1, False, {3: (0.25, 0.75, "outlet")})
    sc = sim.StandardSolverConfig([0, 2], [], [1], False, 1e-8, True, "multigrid", [2, 2])
    a = sim.simple_mesh(mc, sc)
    b = oms.simple_mesh([8, 6], [1.0, 0.75], False, [0, 2], [1], {3: (0.25, 0.75, "outlet")})
    for k in (1, 3, 5, 6, 7, 8, 9):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_set_face_velocity_matches_oracle():
    import jax_fem_cfd.setup.simulation as sim
    s, surf, h, U_d = ns.build([8, 8], [1, 1], False, [0, 1, 2, 3], [], 1e-8, True, "multigrid", [2, 2])
    a = sim.set_face_velocity(U_d, surf[2][1:-1], s.gamma_d, 1.0, s.nn, 1)
    b = ns.cavity_lid(s, surf, U_d)
    assert np.array_equal(a, b)


def test_grid_recognition():
    from jax_fem_cfd._context import infer_grid
    _, nconn, *_ = oms.create_mesh_2d(6, 5, 1, 1)
    assert infer_grid(nconn, 42) == (6, 5, False)
    _, nconn, *_ = oms.create_periodic_mesh_2d(6, 4, 1, 1)
    assert infer_grid(nconn, 24) == (6, 4, True)
    bad = nconn.copy()
    bad[-1] = bad[-1][::-1]
    with pytest.raises(ValueError):
        infer_grid(bad, 24)


def test_multigrid_setup_tuple_matches_reference_layout():
    from jax_fem_cfd.core import multigrid as mg
    from jax_fem_cfd.core import shape_functions as sf
    from jax_fem_cfd.setup import regular_mesh as rm
    lvl, Le_lvls, nconn_lvls, Ld, gp = mg.multigrid_setup([16, 16], [4, 4], [1, 1], [], rm.create_mesh_2d,
                                                           sf.quad4_shape_functions, 1, None)
    assert lvl[0] == (0, 0, 0, 0, 0)
    assert lvl[1] == (289, 17, 17, 256, 1) and lvl[3] == (289 + 81 + 25, 5, 5, 256 + 64 + 16, 3)
    assert Le_lvls.shape == (3, 2, 4, 4) and np.allclose(np.diag(Le_lvls[0, 1]), 0)
    from oracle import multigrid as omg
    from oracle import elements as oel
    ol = omg.multigrid_setup([16, 16], [4, 4], [1, 1], [], oms.create_mesh_2d, oel.quad4_shape_functions)
    assert np.allclose(Ld[:289], ol[0]["L_diag"], rtol=0, atol=1e-15)


def test_host_shape_functions_match_oracle():
    from jax_fem_cfd.core import shape_functions as sf
    from oracle import elements as oel
    c = np.array([[0, 0], [0.1, 0], [0.1, 0.1], [0, 0.1]])
    for a, b in zip(sf.quad4_shape_functions(c), oel.quad4_shape_functions(c)):
        assert np.array_equal(a, b)
    c3 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1.0]]) * 0.5
    for a, b in zip(sf.hex8_shape_functions(c3), oel.hex8_shape_functions(c3)):
        assert np.allclose(a, b, rtol=0, atol=1e-15)
