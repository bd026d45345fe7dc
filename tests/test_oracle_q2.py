"""Pins of the oracle's quad9 (Q2) pieces: mesh structure, closed-form element matrices,
convection matrices against an independent quadrature where the reference's 3x3 Gauss rule is
exact, the lin3 boundary integral and a full step on a flow it must keep."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import elements as el
from oracle import emult
from oracle import mesh as ms
from oracle import ns

RNG = np.random.default_rng(909)
POS = [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (0, 1), (1, 1)]
PX = np.array([p[0] for p in POS])
PY = np.array([p[1] for p in POS])


def test_quad9_mesh_structure():
    nex, ney = 4, 3
    coords, nconn, surf, bconns, nvecs = ms.create_mesh_2d_quad9(nex, ney, 2.0, 1.5)
    nnx, nny = 2 * nex + 1, 2 * ney + 1
    assert coords.shape == (nnx * nny, 2) and nconn.shape == (nex * ney, 9)
    assert np.allclose(coords[nnx + 1], [0.25, 0.25])
    # element e at (ie, je): node at local (a, b) is (2 ie + a, 2 je + b)
    for e in range(nex * ney):
        ie, je = e % nex, e // nex
        assert all(nconn[e, k] == (2 * ie + a) + (2 * je + b) * nnx for k, (a, b) in enumerate(POS))
    cnt = np.bincount(nconn.ravel(), minlength=nnx * nny).reshape(nny, nnx)
    assert cnt[2, 2] == 4 and cnt[1, 2] == 2 and cnt[2, 1] == 2 and cnt[1, 1] == 1 and cnt[0, 0] == 1
    assert np.array_equal(bconns[0][1], [2, 3, 4]) and bconns[3].shape == (ney, 3)
    pc, pn, *_ = ms.create_periodic_mesh_2d_quad9(nex, ney, 2.0, 1.5)
    assert pc.shape == (4 * nex * ney, 2)
    assert np.all(np.bincount(pn.ravel()).reshape(2 * ney, 2 * nex)[::2, ::2] == 4)


# 1D quadratic Lagrange on [0, h], local nodes 0, h/2, h (positions 0, 1, 2)
def _m1(h):
    return h / 30 * np.array([[4.0, 2, -1], [2, 16, 2], [-1, 2, 4]])


def _k1(h):
    return 1 / (3 * h) * np.array([[7.0, -8, 1], [-8, 16, -8], [1, -8, 7]])


_G1 = np.array([[-0.5, 2 / 3, -1 / 6], [-2 / 3, 0.0, 2 / 3], [1 / 6, -2 / 3, 0.5]])  # int N_a N_b'


def _t(ax, ay):
    return ax[np.ix_(PX, PX)] * ay[np.ix_(PY, PY)]


@pytest.mark.parametrize("h", [(0.25, 0.25), (2.0 / 16, 0.5 / 16)])
def test_quad9_closed_form_element_matrices(h):
    hx, hy = h
    coords = np.array([[a * hx / 2, b * hy / 2] for a, b in POS])
    wq, N, Nd = el.quad9_shape_functions(coords)
    Me = el.mass_element(wq, N)
    Le = el.laplacian_element_calc(wq, Nd)
    Ge = el.gradient_element_calc(wq, N, Nd, 2)
    assert np.allclose(Me, _t(_m1(hx), _m1(hy)), rtol=0, atol=1e-16)
    assert np.allclose(Le, _t(_k1(hx), _m1(hy)) + _t(_m1(hx), _k1(hy)), rtol=1e-12, atol=1e-14)
    assert np.allclose(Ge[0], _t(_G1, _m1(hy)), atol=1e-15)
    assert np.allclose(Ge[1], _t(_m1(hx), _G1), atol=1e-15)
    assert np.isclose(Me.sum(), hx * hy)
    wl, Nl, _ = el.lin3_shape_functions(np.array([[0.0, 0.0], [0.0, hy / 2], [0.0, hy]]))
    assert np.allclose((Nl * wl).sum(axis=0), hy / 6 * np.array([1, 4, 1]))


def test_quad9_convection_matrix_brute_force():
    """For a velocity linear in x and y the integrand of Ce has degree <= 5 per direction, so
    the reference's 3x3 Gauss rule is exact; an independent 5x5 rule gives the same matrices."""
    nex, ney, lx, ly = 3, 2, 1.5, 1.0
    coords, nconn, *_ = ms.create_mesh_2d_quad9(nex, ney, lx, ly)
    nn = coords.shape[0]
    x, y = coords[:, 0], coords[:, 1]
    U = np.concatenate([0.3 + 1.1 * x - 0.7 * y, -0.2 + 0.4 * x + 0.9 * y])
    wq, N, Nd = el.quad9_shape_functions(coords[nconn[0]])
    hx, hy = lx / nex, ly / ney
    g, gw = np.polynomial.legendre.leggauss(5)
    L = lambda t: np.array([t * (t - 1) / 2, 1 - t ** 2, t * (t + 1) / 2])
    dL = lambda t: np.array([(2 * t - 1) / 2, -2 * t, (2 * t + 1) / 2])
    for beta in (0.5, 1.0):
        Ce = el.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, beta)
        for e in (0, 4):
            ul = np.stack([U[nconn[e]], U[nconn[e] + nn]], axis=1)
            ref = np.zeros((9, 9))
            for a, wa in zip(g, gw):
                for b, wb in zip(g, gw):
                    Nq = L(a)[PX] * L(b)[PY]
                    dNx = dL(a)[PX] * L(b)[PY] * (2 / hx)
                    dNy = L(a)[PX] * dL(b)[PY] * (2 / hy)
                    uq = Nq @ ul
                    div = dNx @ ul[:, 0] + dNy @ ul[:, 1]
                    ref += wa * wb * hx * hy / 4 * (np.outer(Nq, uq[0] * dNx + uq[1] * dNy)
                                                    + beta * div * np.outer(Nq, Nq))
            assert np.allclose(Ce[e], ref, atol=1e-14)


def test_quad9_boundary_integral_inflow():
    """Uniform inflow u = 1 on the left face: Simpson shares h/6 (1, 4, 1) per edge element."""
    n, L = 4, 1.0
    s, surf, h, U_d = ns.build([n, n], [L, L], False, [3], [], 1e-8, True, None, None, shape_func_ord=2)
    U_d = ms.set_face_velocity(U_d, surf[3], s.gamma_d, 1.0, s.nn, 1)
    F_p = el.u_dot_n_boundary_integral_1d(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords,
                                          el.lin3_shape_functions)
    he = L / n
    ref = np.zeros(2 * n + 1)
    for k in range(n):
        ref[2 * k:2 * k + 3] += he / 6 * np.array([1, 4, 1])
    assert np.allclose(F_p[surf[3]], -ref) and np.isclose(F_p.sum(), -L)


def test_quad9_matvecs_match_assembly():
    coords, nconn, *_ = ms.create_mesh_2d_quad9(3, 2, 1.5, 1.0)
    nn = coords.shape[0]
    wq, N, Nd = el.quad9_shape_functions(coords[nconn[0]])
    Le = el.laplacian_element_calc(wq, Nd)
    rows = np.repeat(nconn, 9, axis=1).ravel()
    cols = np.tile(nconn, (1, 9)).ravel()
    A = sp.coo_matrix((np.tile(Le.ravel(), nconn.shape[0]), (rows, cols)), shape=(nn, nn)).tocsr()
    P = RNG.standard_normal(nn)
    assert np.allclose(emult.laplacian_element_mult(P, Le, nconn, nn), A @ P, atol=1e-13)
    assert np.allclose(A @ np.ones(nn), 0, atol=1e-13)


@pytest.mark.parametrize("precond", ["jacobi", "multigrid"])
def test_quad9_step_keeps_uniform_flow(precond):
    n, L = [8, 8], [1.0, 1.0]
    s, surf, h, U_d = ns.build(n, L, False, [0, 1, 2, 3], [], 1e-12, True, precond, [2, 2], shape_func_ord=2)
    g = s.gamma_d[s.gamma_d < s.nn]
    U_d = ms.set_face_velocity(U_d, g, s.gamma_d, 1.0, s.nn, 1)
    U0 = np.zeros(2 * s.nn)
    U0[:s.nn] = 1.0
    F_p, _ = ns.compute(s, U_d, None, 0, 0)
    U, P, ui, pi = ns.four_step_timestep(s, U0, np.zeros(s.nn), 1e-2, 1.0, 1e-2, U_d, F_p, 0)
    assert np.allclose(U, U0, atol=1e-11) and np.allclose(P, 0, atol=1e-11)


def test_quad9_taylor_green_converges():
    """Periodic Taylor-Green, 20 steps of dt = 1e-3: the quad9 velocity error falls at least
    at second order between 8x8 and 16x16 elements.  (With the row-sum lumped mass of the
    projection, quad9 corner / mid-side / centre weights are 1/36, 1/9, 4/9 of the element area,
    so the constant is larger than Q1's on the same grid; the order is what pins the path.)"""
    errs = []
    for n in (8, 16):
        s, surf, h, U_d = ns.build([n, n], [2 * np.pi, 2 * np.pi], True, [], [], 1e-12, False, "multigrid",
                                   [2, 2], shape_func_ord=2)
        u, v, p = ns.taylor_green(s.coords, 0.0, 1.0, 0.01)
        U, P = np.concatenate([u, v]), p
        for _ in range(20):
            U, P, _, _ = ns.four_step_timestep(s, U, P, 1e-3, 1.0, 0.01, None, 0, 0)
        ua, va, _ = ns.taylor_green(s.coords, 0.02, 1.0, 0.01)
        errs.append(ns.rel_l2(U, np.concatenate([ua, va])))
    assert errs[1] < 0.25 * errs[0] and errs[1] < 2e-3, errs
