"""GPU parity of the reference's multigrid building blocks as standalone host operators
(jax_fem_cfd/core/multigrid.py:90-228), the module-level solver functions of
jax_fem_cfd/solvers/standard.py and the BiCGSTAB breakdown codes of the device solver
(jax_fem_cfd/utils/jax_iterative_solvers.py:156-157), all through the C ABI, against the oracle.

Tolerances: transfers and the Jacobi sweep are exact up to summation order (1e-14 relative);
solves 1e-10 with equal iteration counts; breakdown codes exact.
"""
import numpy as np
import pytest

from oracle import emult as oem
from oracle import multigrid as omg
from oracle import mesh as oms
from oracle import elements as oel
from oracle import ns

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(7)


def _np(x):
    return x.detach().cpu().numpy()


@pytest.mark.parametrize("periodic", [False, True])
def test_interpolation_and_restriction_operators(periodic):
    """multigrid.py:90-212 with the reference's own arguments: coarse 9x7 elements (10x8 nodes,
    the periodic call counts the padded wrap column and row), fine 18x14."""
    from jax_fem_cfd.core import multigrid as mg
    nHx, nHy = 10, 8
    nhx, nhy = 2 * (nHx - 1) + 1, 2 * (nHy - 1) + 1
    nH = (nHx - 1) * (nHy - 1) if periodic else nHx * nHy
    nh = (nhx - 1) * (nhy - 1) if periodic else nhx * nhy
    print(f"\n[case] transfers periodic={periodic}", flush=True)
    xH = RNG.standard_normal(nH)
    xh = RNG.standard_normal(nh)
    got_i = _np(mg.interpolation_operator(xH, nHx, nHy, nhx, nhy, periodic))
    ref_i = omg.interpolation_operator(xH, nHx, nHy, nhx, nhy, periodic)
    got_r = _np(mg.restriction_operator(xh, nhx, nhy, nHx, nHy, periodic))
    ref_r = omg.restriction_operator(xh, nhx, nhy, nHx, nHy, periodic)
    assert np.allclose(got_i, ref_i, rtol=1e-14, atol=1e-15)
    assert np.allclose(got_r, ref_r, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_single_jacobi_iter(set_zeroP):
    """multigrid.py:214-228 on a 12x10 grid with a nonzero initial x (pins zeroed before and
    after, or the mean subtracted)."""
    from jax_fem_cfd.core import multigrid as mg
    coords, nconn, *_ = oms.create_mesh_2d(12, 10, 1.2, 1.0)
    nn = coords.shape[0]
    wq, N, Nd = oel.quad4_shape_functions(coords[nconn[0]])
    Le = oel.laplacian_element_calc(wq, Nd)
    Le_off = Le - np.diag(np.diag(Le))
    Ld = oel.laplacian_diag(Le, nconn, nn)
    gamma_p = np.array([0, 5, 40])
    x = RNG.standard_normal(nn)
    rhs = RNG.standard_normal(nn)
    print(f"\n[case] Jacobi sweep set_zeroP={set_zeroP}", flush=True)
    got = _np(mg.single_jacobi_iter(x, rhs, Le_off, Ld, gamma_p, nconn, nn, set_zeroP))
    ref = omg.single_jacobi_iter(x.copy(), rhs, Le_off, Ld, gamma_p, nconn, nn, set_zeroP)
    bad = np.flatnonzero(~np.isclose(got, ref, rtol=1e-14, atol=1e-14))
    print(f"mismatches {bad.size}: nodes {bad[:8]} got {got[bad[:4]]} ref {ref[bad[:4]]}", flush=True)
    assert bad.size == 0
    if set_zeroP:
        assert np.all(got[gamma_p] == 0.0)
    else:
        assert abs(got.mean()) < 1e-15 * np.abs(got).max() * nn


def test_standard_module_functions():
    """solve_momentum / solve_pressure / four_step_timestep / compute as module-level functions
    with the reference's argument lists (standard.py:14, :85, :148, :168) on a 32x32 cavity."""
    from jax_fem_cfd.solvers import standard
    n = 32
    s, surf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, "multigrid", [4, 4])
    oUd = ns.cavity_lid(s, surf, oUd)
    U0 = np.random.default_rng(1).normal(0, 1e-3, 2 * s.nn)
    U0[s.gamma_d] = oUd
    P0 = RNG.standard_normal(s.nn) * 0.1
    P0[0] = 0.0
    print("\n[case] standard.* module functions", flush=True)
    Ce = oel.conservative_convec_elem_calc(s.nconn, U0, s.wq, s.N, s.Nderiv, s.nn, 2, 0.5)
    ref_U, ref_k = ns.solve_momentum(U0, P0, s.Me, s.Le, s.Ge, Ce, 1.0, 1e-3, 5e-4, s.nconn, s.nn, 2, s.gamma_d,
                                     oUd, 0, 1e-10, 0.5, 1)
    got_U, got_k = standard.solve_momentum(U0, P0, s.Me, s.Le, s.Ge, Ce, 1.0, 1e-3, 5e-4, s.nconn, s.nn, 2,
                                           s.gamma_d, oUd, 0, 1e-10, 0.5, 1)
    assert got_k == ref_k and ns.rel_l2(_np(got_U), ref_U) < 1e-10
    Ustar = ref_U
    ref_P, ref_pk = ns.solve_pressure(Ustar, P0, s.Le, s.Ge, 1.0, 5e-4, s.nconn, s.nn, 2, s.gamma_p, 0, True,
                                      s.M_op, 1e-10, 0)
    got_P, got_pk = standard.solve_pressure(Ustar, P0, s.Le, s.Ge, 1.0, 5e-4, s.nconn, s.nn, 2, s.gamma_p, 0,
                                            True, "multigrid", [4, 4], 1e-10, 0)
    assert got_pk == ref_pk and ns.rel_l2(_np(got_P), ref_P) < 1e-10
    # the whole step and compute() with the reference's argument lists
    Uo, Po, uio, pio = ns.four_step_timestep(s, U0.copy(), P0.copy(), 5e-4, 1.0, 1e-3, oUd, 0, 0)
    U, P, ui, pi = standard.four_step_timestep(U0, P0, s.M, s.Me, s.Le, s.Ge, 1.0, 1e-3, 5e-4, s.nconn, s.nn, 2,
                                               s.wq, s.N, s.Nderiv, s.gamma_d, oUd, s.gamma_p, 0, 0, True,
                                               "multigrid", [4, 4], 1e-10, 0.5)
    assert (ui, pi) == (uio, pio)
    assert ns.rel_l2(_np(U), Uo) < 1e-10 and ns.rel_l2(_np(P), Po) < 1e-10


def test_passive_update_phi_with_reference_arguments():
    """passive_transport.update_phi with its argument list (convection matrices given) against
    the oracle's update_phi on a 24x24 walled box."""
    from jax_fem_cfd.solvers import passive_transport as pt
    n = 24
    s, surf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, "multigrid", [4, 4])
    U_n = RNG.normal(0, 0.3, 2 * s.nn)
    U = U_n + RNG.normal(0, 0.05, 2 * s.nn)
    phi_n = RNG.random(s.nn)
    Ce_n = oel.conservative_convec_elem_calc(s.nconn, U_n, s.wq, s.N, s.Nderiv, s.nn, 2, 1.0)
    Ce = oel.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 2, 1.0)
    gd = np.array(surf[0])
    phi_d = np.linspace(0.0, 1.0, len(gd))
    print("\n[case] update_phi with the reference's arguments", flush=True)
    ref, kr = ns.update_phi(phi_n, s.nconn, s.Me, Ce, Ce_n, s.Le, 1e-2, 1e-3, s.nn, gd, phi_d, 1e-10)
    got, k = pt.update_phi(phi_n, s.nconn, s.Me, Ce, Ce_n, s.Le, 1e-2, 1e-3, s.nn, gd, phi_d, 1e-10)
    assert abs(k - kr) <= 1 and ns.rel_l2(_np(got), ref) < 1e-9


@pytest.mark.parametrize("case", ["omega", "rho"])
def test_bicgstab_breakdown_codes(case):
    """utils.jax_iterative_solvers.bicgstab on device vectors returns the reference's breakdown
    codes: -11 when omega = t.s / t.t is 0 (2x2 case), -10 when rhat.r is 0 in the second step
    (3x3 lower-bidiagonal case); x equals the oracle's."""
    from oracle.solvers import bicgstab as obicg
    from jax_fem_cfd.utils.jax_iterative_solvers import bicgstab
    from jax_fem_cfd import _native as nat
    t = nat.torch()
    if case == "omega":
        A = np.array([[-2.0, -2.0], [-2.0, 0.0]])
        b, code = np.array([1.0, 0.0]), -11
    else:
        A = np.array([[1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 1.0]])
        b, code = np.array([1.0, 0.0, 0.0]), -10
    Ad = t.tensor(A, dtype=t.float64, device=nat.device())
    print(f"\n[case] BiCGSTAB breakdown {case}", flush=True)
    x, k = bicgstab(lambda v: Ad @ v, b, tol=1e-12, maxiter=50)
    with np.errstate(all="ignore"):
        xr, kr = obicg(lambda v: A @ v, b, tol=1e-12, maxiter=50)
    assert k == kr == code
    assert np.array_equal(_np(x), xr, equal_nan=True)  # alpha = 0/0 in the rho case, as in the reference
