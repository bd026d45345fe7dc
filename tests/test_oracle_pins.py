"""Pins the CPU oracle against what the mathematics fixes: closed forms,
invariants, independent scipy assemblies/solves and analytic solutions.
Each check is chosen so that a plausible slip (dropped term, wrong sign,
transposed operand, wrong index) fails it."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import elements as el
from oracle import emult
from oracle import mesh as ms
from oracle import multigrid as mg
from oracle import ns
from oracle.solvers import bicgstab, cg

RNG = np.random.default_rng(1234)


# ----------------------------------------------------------------------------- mesh
@pytest.mark.parametrize("nex,ney", [(3, 2), (8, 8), (5, 7)])
def test_mesh_2d_formulas(nex, ney):
    coords, nconn, surf, bconns, nvecs = ms.create_mesh_2d(nex, ney, 2.0, 3.0)
    nnx = nex + 1
    for e in range(nex * ney):
        ie, je = e % nex, e // nex
        n0 = je * nnx + ie
        assert list(nconn[e]) == [n0, n0 + 1, n0 + nnx + 1, n0 + nnx]
    assert coords.shape == ((nex + 1) * (ney + 1), 2)
    assert np.allclose(coords[nnx + 1], [2.0 / nex, 3.0 / ney])
    counts = np.bincount(nconn.ravel())
    assert counts.sum() == 4 * nex * ney
    interior = counts.reshape(ney + 1, nex + 1)[1:-1, 1:-1]
    assert np.all(interior == 4)
    assert list(surf[0]) == list(range(nnx))
    assert surf[1][-1] == (ney + 1) * nnx - 1 and surf[3][1] == nnx
    assert np.all(bconns[2][:, 1] - bconns[2][:, 0] == 1)
    assert np.all(nvecs[3] == [-1, 0])


def test_periodic_mesh_wraps():
    _, nconn, surf, _, _ = ms.create_periodic_mesh_2d(4, 3, 1.0, 1.0)
    assert surf is None
    assert list(nconn[3]) == [3, 0, 4, 7]          # last column wraps in x
    assert list(nconn[11]) == [11, 8, 0, 3]        # top-right wraps in x and y
    assert np.all(np.bincount(nconn.ravel()) == 4)


def test_bc_index_sets_cavity():
    s, surf, h, U_d = ns.build([8, 8], [1, 1], False, [0, 1, 2, 3], [], 1e-8, True, "multigrid", [2, 2])
    nn = s.nn
    g = s.gamma_d
    assert len(g) == 2 * 32 and np.all(np.diff(g[:32]) > 0)
    assert np.all(g[32:] == g[:32] + nn)
    assert s.gamma_p == 0
    Ud = ns.cavity_lid(s, surf, U_d)
    lid = surf[2][1:-1]
    assert np.all(Ud[np.searchsorted(g, lid)] == 1.0) and Ud.sum() == len(lid)


# ----------------------------------------------------------------------------- element matrices
@pytest.mark.parametrize("h", [1.0 / 16, 2 * np.pi / 64])
def test_closed_form_element_matrices(h):
    coords = np.array([[0, 0], [h, 0], [h, h], [0, h]], dtype=float)
    wq, N, Nd = el.quad4_shape_functions(coords)
    Me = el.mass_element(wq, N)
    Le = el.laplacian_element_calc(wq, Nd)
    Ge = el.gradient_element_calc(wq, N, Nd, 2)
    Me_cf = h * h / 36 * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]])
    Le_cf = np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]]) / 6
    Gx = h / 12 * np.array([[-2, 2, 1, -1], [-2, 2, 1, -1], [-1, 1, 2, -2], [-1, 1, 2, -2]])
    Gy = h / 12 * np.array([[-2, -1, 1, 2], [-1, -2, 2, 1], [-1, -2, 2, 1], [-2, -1, 1, 2]])
    assert np.allclose(Me, Me_cf, rtol=0, atol=1e-15 * max(1, h * h))
    assert np.allclose(Le, Le_cf, rtol=0, atol=1e-14)
    assert np.allclose(Ge[0], Gx, atol=1e-15) and np.allclose(Ge[1], Gy, atol=1e-15)
    assert np.allclose(Le @ np.ones(4), 0, atol=1e-14)
    assert np.allclose(Ge @ np.ones(4), 0, atol=1e-15)


def test_lumped_mass_and_diag():
    coords, nconn, *_ = ms.create_mesh_2d(4, 4, 1.0, 1.0)
    h = 0.25
    wq, N, Nd = el.quad4_shape_functions(coords[nconn[0]])
    M, _ = el.lumped_mass_matrix(nconn, wq, N, coords.shape[0], 2)
    M = M[:25].reshape(5, 5)
    assert np.isclose(M[2, 2], h * h) and np.isclose(M[0, 2], h * h / 2) and np.isclose(M[0, 0], h * h / 4)
    Ld = el.laplacian_diag(el.laplacian_element_calc(wq, Nd), nconn, 25).reshape(5, 5)
    assert np.isclose(Ld[2, 2], 8 / 3) and np.isclose(Ld[0, 2], 4 / 3) and np.isclose(Ld[0, 0], 2 / 3)


def _gauss(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def test_convection_matrix_brute_force_quadrature():
    """Ce from a 5x5 Gauss rule written independently (bilinear u, h-scaled
    derivatives) equals the 2x2-point oracle: 2x2 is exact for the Q1
    integrand, so a wrong index or sign shows up."""
    nex = 3
    coords, nconn, *_ = ms.create_mesh_2d(nex, nex, 1.0, 1.0)
    nn = coords.shape[0]
    h = 1.0 / nex
    U = RNG.standard_normal(2 * nn)
    wq, N, Nd = el.quad4_shape_functions(coords[nconn[0]])
    for beta in (0.0, 0.5, 1.0):
        Ce = el.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, beta)
        x, w = _gauss(5)
        for e in (0, 4, 8):
            ul = np.stack([U[nconn[e]], U[nconn[e] + nn]], axis=1)
            ref = np.zeros((4, 4))
            for xi, wx in zip(x, w):
                for et, we in zip(x, w):
                    sx = np.array([-1, 1, 1, -1.0]); sy = np.array([-1, -1, 1, 1.0])
                    Nq = (1 + sx * xi) * (1 + sy * et) / 4
                    dNx = sx * (1 + sy * et) / 4 * (2 / h)
                    dNy = (1 + sx * xi) * sy / 4 * (2 / h)
                    uq = Nq @ ul
                    div = dNx @ ul[:, 0] + dNy @ ul[:, 1]
                    jac = wx * we * h * h / 4
                    ref += jac * (np.outer(Nq, uq[0] * dNx + uq[1] * dNy) + beta * div * np.outer(Nq, Nq))
            assert np.allclose(Ce[e], ref, atol=1e-14)


def test_convection_skew_symmetric_periodic():
    """beta = 1/2 on a periodic mesh: the assembled C is skew-symmetric, so
    u^T C u = 0 for any u."""
    coords, nconn, *_ = ms.create_periodic_mesh_2d(8, 8, 2 * np.pi, 2 * np.pi)
    nn = coords.shape[0]
    U = RNG.standard_normal(2 * nn)
    wq, N, Nd = el.quad4_shape_functions(coords[nconn[0]])
    C = emult.assemble(nconn, nn, el.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, 0.5))
    assert abs((C + C.T)).max() < 1e-13
    C0 = emult.assemble(nconn, nn, el.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, 0.0))
    assert abs((C0 + C0.T)).max() > 1e-3


# ----------------------------------------------------------------------------- matvecs
def _setup_small(periodic=False, n=6):
    if periodic:
        coords, nconn, *_ = ms.create_periodic_mesh_2d(n, n, 1.0, 1.0)
    else:
        coords, nconn, *_ = ms.create_mesh_2d(n, n + 1, 1.0, 1.2)
    nn = coords.shape[0]
    wq, N, Nd = el.quad4_shape_functions(coords[nconn[0]])
    return coords, nconn, nn, wq, N, Nd


@pytest.mark.parametrize("periodic", [False, True])
def test_matvecs_match_scipy_assembly(periodic):
    coords, nconn, nn, wq, N, Nd = _setup_small(periodic)
    Me = el.mass_element(wq, N)
    Le = el.laplacian_element_calc(wq, Nd)
    Ge = el.gradient_element_calc(wq, N, Nd, 2)
    U = RNG.standard_normal(2 * nn)
    P = RNG.standard_normal(nn)
    Ce = el.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, 0.5)
    A_M, A_L = emult.assemble(nconn, nn, Me), emult.assemble(nconn, nn, Le)
    A_G = [emult.assemble(nconn, nn, Ge[i]) for i in range(2)]
    A_C = emult.assemble(nconn, nn, Ce)
    u, v = U[:nn], U[nn:]
    assert np.allclose(emult.consistent_mass_mult(U, Me, nconn, nn, 2), np.concatenate([A_M @ u, A_M @ v]))
    assert np.allclose(emult.laplacian_element_mult(P, Le, nconn, nn), A_L @ P)
    assert np.allclose(emult.gradient_element_mult(P, Ge, nconn, nn, 2), np.concatenate([A_G[0] @ P, A_G[1] @ P]))
    assert np.allclose(emult.divergence_element_mult(U, Ge, nconn, nn, 2), A_G[0].T @ u + A_G[1].T @ v)
    rho, mu = 1.3, 0.07
    K = rho * A_C + mu * A_L
    assert np.allclose(emult.visc_conv_mult(U, Ce, Le, mu, rho, nconn, nn, 2), np.concatenate([K @ u, K @ v]))
    # adjoint identity <G p, u> = <p, D u>
    Gp = emult.gradient_element_mult(P, Ge, nconn, nn, 2)
    assert np.isclose(Gp @ U, P @ emult.divergence_element_mult(U, Ge, nconn, nn, 2))


def test_nine_point_stencil():
    coords, nconn, *_ = ms.create_mesh_2d(6, 6, 1.0, 1.0)
    wq, N, Nd = el.quad4_shape_functions(coords[nconn[0]])
    L = emult.assemble(nconn, 49, el.laplacian_element_calc(wq, Nd)).toarray()
    c = 3 * 7 + 3
    row = L[c].reshape(7, 7)[2:5, 2:5]
    assert np.allclose(row, np.array([[-1, -1, -1], [-1, 8, -1], [-1, -1, -1]]) / 3)


# ----------------------------------------------------------------------------- solvers
def test_cg_and_bicgstab_against_spsolve():
    coords, nconn, nn, wq, N, Nd = _setup_small(False, 10)
    Le = el.laplacian_element_calc(wq, Nd)
    Me = el.mass_element(wq, N)
    A = (emult.assemble(nconn, nn, Le) + 50 * emult.assemble(nconn, nn, Me)).tocsr()
    b = RNG.standard_normal(nn)
    xs = spla.spsolve(A.tocsc(), b)
    x, k = cg(lambda v: A @ v, b, tol=1e-12)
    assert np.linalg.norm(b - A @ x) <= 1e-12 * np.linalg.norm(b) * 1.0001
    assert np.allclose(x, xs, rtol=1e-9, atol=1e-11) and k > 3
    # scipy's cg runs the same recurrence: same iteration count
    cnt = [0]
    spla.cg(A, b, rtol=1e-12, atol=0.0, callback=lambda _: cnt.__setitem__(0, cnt[0] + 1))
    assert abs(cnt[0] - k) <= 1
    Ce = el.conservative_convec_elem_calc(nconn, RNG.standard_normal(2 * nn) * 5, wq, N, Nd, nn, 2, 0.5)
    B = (A + emult.assemble(nconn, nn, Ce)).tocsr()
    x, k = bicgstab(lambda v: B @ v, b, tol=1e-12)
    assert k > 0 and np.linalg.norm(b - B @ x) <= 2e-12 * np.linalg.norm(b)
    assert np.allclose(x, spla.spsolve(B.tocsc(), b), rtol=1e-8, atol=1e-10)


def test_bicgstab_breakdown_codes_and_maxiter():
    A = lambda v: v  # noqa: E731
    x, k = bicgstab(A, np.zeros(5), tol=1e-8)
    assert k == 0 and np.all(x == 0)
    x, k = cg(lambda v: 2 * v + np.roll(v, 1), np.ones(7), tol=1e-14, maxiter=1)
    assert k == 1
    # omega = t.s / t.t = 0 in the first step: s = (0, 1), t = A s = (-2, 0)
    A2 = np.array([[-2.0, -2.0], [-2.0, 0.0]])
    with np.errstate(all="ignore"):
        x, k = bicgstab(lambda v: A2 @ v, np.array([1.0, 0.0]), tol=1e-12, maxiter=50)
    assert k == -11
    # rho breakdown in the second step: b = e1 gives alpha = 1, s = (0, -1, 0), t = (0, -1, -1),
    # omega = 1/2, r1 = (0, -1/2, 1/2) with rhat.r1 = 0 exactly while r1.r1 = 1/2
    A3 = np.array([[1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 1.0]])
    with np.errstate(all="ignore"):
        x, k = bicgstab(lambda v: A3 @ v, np.array([1.0, 0.0, 0.0]), tol=1e-12, maxiter=50)
    assert k == -10


# ----------------------------------------------------------------------------- multigrid
def _dense(op, n):
    return np.stack([op(np.eye(n)[i]) for i in range(n)], axis=1)


def test_restriction_is_transpose_of_interpolation_nonperiodic():
    nnHx, nnHy, nnhx, nnhy = 5, 4, 9, 7
    I = _dense(lambda v: mg.interpolation_operator(v, nnHx, nnHy, nnhx, nnhy, False), nnHx * nnHy)
    R = _dense(lambda v: mg.restriction_operator(v, nnhx, nnhy, nnHx, nnHy, False), nnhx * nnhy)
    assert np.allclose(R, I.T, atol=0)
    assert np.allclose(I @ np.ones(nnHx * nnHy), 1.0)
    R1 = (R @ np.ones(nnhx * nnhy)).reshape(nnHy, nnHx)
    assert R1[1, 1] == 4 and R1[0, 1] == 3 and R1[1, 0] == 3 and R1[0, 0] == 2.25 and R1[-1, -1] == 2.25


def test_periodic_transfer_quirk():
    """Periodic prolongation wraps; the restriction drops the wrapped
    contributions (reference crops the padded coarse grid)."""
    nH, nh = 4, 8
    I = _dense(lambda v: mg.interpolation_operator(v, nH + 1, nH + 1, nh + 1, nh + 1, True), nH * nH)
    R = _dense(lambda v: mg.restriction_operator(v, nh + 1, nh + 1, nH + 1, nH + 1, True), nh * nh)
    assert np.allclose(I @ np.ones(nH * nH), 1.0)
    R1 = (R @ np.ones(nh * nh)).reshape(nH, nH)
    assert R1[0, 0] == 2.25 and R1[0, 1] == 3 and R1[1, 0] == 3 and R1[2, 3] == 4
    assert not np.allclose(R, I.T)


def test_vcycle_neumann_is_linear():
    s, *_ = ns.build([16, 16], [1, 1], False, [], [], 1e-8, False, "multigrid", [2, 2])
    r1, r2 = RNG.standard_normal(s.nn), RNG.standard_normal(s.nn)
    a, b = 0.7, -1.9
    lhs = s.M_op(a * r1 + b * r2)
    rhs = a * s.M_op(r1) + b * s.M_op(r2)
    assert np.allclose(lhs, rhs, atol=1e-12)
    assert abs(s.M_op(r1).mean()) < 1e-14


def test_two_level_vcycle_dense_restatement():
    """Two-level pinned V-cycle written with dense matrices (Jacobi from
    zero, residual, R, exact coarse solve, I, Jacobi) equals the oracle's
    loop-based V-cycle to the coarse CG tolerance."""
    s, *_ = ns.build([4, 4], [1, 1], False, [0, 1, 2, 3], [], 1e-14, True, "multigrid", [2, 2])
    f, c = s.levels
    Lf = emult.assemble(f["nconn"], f["nn"], f["Le"]).toarray()
    Lc = emult.assemble(c["nconn"], c["nn"], c["Le"]).toarray()
    I = _dense(lambda v: mg.interpolation_operator(v, 3, 3, 5, 5, False), 9)
    R = I.T
    Df = np.diag(Lf)
    r = RNG.standard_normal(25); r[0] = 0
    Pf = np.eye(25); Pf[0, 0] = 0
    x1 = Pf @ (2 / 3 * r / Df)
    rH = R @ (r - Pf @ (Lf @ x1))
    Lcp = Lc.copy(); Lcp[0, :] = 0; Lcp[0, 0] = 1
    rHp = rH.copy(); rHp[0] = 0
    xH = np.linalg.solve(Lcp, rHp); xH[0] = 0
    x = x1 + I @ xH
    x[0] = 0
    off = (Lf - np.diag(Df)) @ x
    x = 2 / 3 * (r - off) / Df + 1 / 3 * x
    x[0] = 0
    assert np.allclose(s.M_op(r), x, atol=1e-10)


# ----------------------------------------------------------------------------- Navier-Stokes
def test_taylor_green_tracks_analytic():
    s, surf, h, U_d = ns.build([32, 32], [2 * np.pi, 2 * np.pi], True, [], [], 1e-10, False, "multigrid", [2, 2])
    rho, mu, dt = 1.0, 0.01, 1e-3
    u, v, p = ns.taylor_green(s.coords, 0, rho, mu)
    U, P, ui, pi, *_ = ns.timestep_loop(s, 20 * dt, dt, 10, rho, mu, U_d, np.concatenate([u, v]), p - p.mean())
    ue, ve, pe = ns.taylor_green(s.coords, 20 * dt, rho, mu)
    assert ns.rel_l2(U, np.concatenate([ue, ve])) < 5e-5
    assert ns.rel_l2(P, pe - pe.mean()) < 5e-3


def test_projection_reduces_divergence_and_pins():
    s, surf, h, U_d = ns.build([16, 16], [1, 1], False, [0, 1, 2, 3], [], 1e-10, True, "multigrid", [4, 4])
    U_d = ns.cavity_lid(s, surf, U_d)
    U0 = np.zeros(2 * s.nn); U0[s.gamma_d] = U_d
    U, P, ui, pi = ns.four_step_timestep(s, U0, np.zeros(s.nn), 1e-3, 1.0, 1e-2, U_d, 0, 0)
    assert P[0] == 0 and np.all(U[s.gamma_d] == U_d)
    F_p, _ = ns.compute(s, U_d, None, 0, 0)
    assert np.allclose(F_p, 0)
    assert ui > 0 and pi > 0


def test_boundary_integral_inflow():
    """Uniform inflow u = (1, 0) on the left face (n = (-1, 0)): F_p = -h/2
    at the face ends and -h at interior face nodes."""
    s, surf, h, U_d = ns.build([4, 4], [1, 1], False, [3], [], 1e-8, True, None, None)
    U_d = ms.set_face_velocity(U_d, surf[3], s.gamma_d, 1.0, s.nn, 1)
    F_p = el.u_dot_n_boundary_integral_1d(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords)
    left = F_p[surf[3]]
    assert np.allclose(left, [-h / 2, -h, -h, -h, -h / 2])
    assert np.isclose(F_p.sum(), -1.0)


def test_passive_approx_inv_variant():
    """The passive solver's approx_inv closure (passive_transport.py:154-162) keeps D1x and D1L on
    the gamma_p rows: there z = r / L_diag^2 exactly, and away from the pins (where neither L D1x
    nor L D1L reads a pinned entry) it equals the standard closure (standard.py:225-234), whose
    pinned rows give z = 3 r / L_diag."""
    n = 8
    s0, *_ = ns.build([n, n], [1.0, 1.0], False, [], [], 1e-10, True, "approx_inv", None)
    s1, *_ = ns.build([n, n], [1.0, 1.0], False, [], [], 1e-10, True, "approx_inv", None, passive=True)
    r = RNG.standard_normal(s0.nn)
    z0, z1 = s0.M_op(r), s1.M_op(r)
    gp = np.atleast_1d(s0.gamma_p)
    Ld = s0.L_diag
    assert np.allclose(z1[gp], r[gp] / Ld[gp] ** 2, rtol=1e-14)
    assert np.allclose(z0[gp], 3 * r[gp] / Ld[gp], rtol=1e-14)
    nnx = n + 1
    ii, jj = np.arange(s0.nn) % nnx, np.arange(s0.nn) // nnx
    far = (ii > 2) | (jj > 2)  # pin is node 0; L D1L reaches two nodes away
    assert np.array_equal(z0[far], z1[far]) and not np.allclose(z0, z1)
