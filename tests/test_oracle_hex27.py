"""Pins of the oracle's triquadratic (hex27) pieces.  The reference has no 3D quadratic element
(simulation.py:73-74 rejects it), so every check here is fixed by the mathematics: the mesh
against its closed-form index structure, the element constants against tensor products of
1D quadratic-element integrals computed by an independent Gauss rule, Ce against an independent
5x5x5 rule, the quad9 face integral against Simpson shares, assembly, and a full step that must
leave a uniform flow unchanged."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import elements as el
from oracle import emult
from oracle import mesh as ms
from oracle import ns

RNG = np.random.default_rng(2727)

# 1D quadratic Lagrange basis on [0, 1] with nodes 0, 1/2, 1 (independent of the oracle's form)
_B1 = [lambda x: 2 * (x - 0.5) * (x - 1), lambda x: -4 * x * (x - 1), lambda x: 2 * x * (x - 0.5)]
_D1 = [lambda x: 4 * x - 3, lambda x: -8 * x + 4, lambda x: 4 * x - 1]


def _one_d(h):
    """1D mass, stiffness and int N_a dN_b/dx on an element of length h (6-point Gauss)."""
    x, w = np.polynomial.legendre.leggauss(6)
    x, w = (x + 1) / 2, w / 2
    M = np.array([[np.sum(w * _B1[a](x) * _B1[b](x)) * h for b in range(3)] for a in range(3)])
    K = np.array([[np.sum(w * _D1[a](x) * _D1[b](x)) / h for b in range(3)] for a in range(3)])
    G = np.array([[np.sum(w * _B1[a](x) * _D1[b](x)) for b in range(3)] for a in range(3)])
    return M, K, G


_POS = [(a, b, c) for c in range(3) for b in range(3) for a in range(3)]
_IX = np.array([p[0] for p in _POS])
_IY = np.array([p[1] for p in _POS])
_IZ = np.array([p[2] for p in _POS])


def _tensor(ax, ay, az):
    return ax[np.ix_(_IX, _IX)] * ay[np.ix_(_IY, _IY)] * az[np.ix_(_IZ, _IZ)]


def test_hex27_mesh_structure():
    nex, ney, nez, L = 3, 2, 2, (1.5, 1.0, 0.5)
    coords, nconn, surf, bconns, nvecs = ms.create_mesh_3d_hex27(nex, ney, nez, *L)
    nnx, nny, nnz = 2 * nex + 1, 2 * ney + 1, 2 * nez + 1
    assert coords.shape == (nnx * nny * nnz, 3) and nconn.shape == (nex * ney * nez, 27)
    h = np.array(L) / np.array([nex, ney, nez])
    for e in range(nconn.shape[0]):
        ke, r = divmod(e, nex * ney)
        je, ie = divmod(r, nex)
        for b, (a, by, c) in enumerate(_POS):
            assert np.allclose(coords[nconn[e, b]], [(2 * ie + a) * h[0] / 2, (2 * je + by) * h[1] / 2,
                                                     (2 * ke + c) * h[2] / 2])
    # incidence: a node with (i, j, k) has prod over axes (2 if even interior index else 1) elements
    cnt = np.bincount(nconn.ravel(), minlength=coords.shape[0]).reshape(nnz, nny, nnx)
    for k in range(nnz):
        for j in range(nny):
            for i in range(nnx):
                f = lambda t, n: 2 if (t % 2 == 0 and 0 < t < n - 1) else 1
                assert cnt[k, j, i] == f(i, nnx) * f(j, nny) * f(k, nnz)
    # faces: the nodes of the bounding planes, face elements in quad9 order, normals outward
    planes = [(1, 0.0), (0, L[0]), (1, L[1]), (0, 0.0), (2, 0.0), (2, L[2])]
    for f, (ax, v) in enumerate(planes):
        on = np.flatnonzero(np.isclose(coords[:, ax], v))
        assert np.array_equal(np.sort(surf[f]), on)
        assert np.all(np.isclose(coords[bconns[f].ravel(), ax], v))
        assert bconns[f].shape[1] == 9 and np.array_equal(np.unique(bconns[f]), on)
        assert np.all(nvecs[f][:, ax] == (1 if v > 0 else -1))
    # x-faces are counterclockwise in (z, y): corner 0 -> 1 along +z, 1 -> 2 along +y
    for f in (1, 3):
        P = coords[bconns[f]][:, :, [2, 1]]
        assert np.all(P[:, 2, 0] > P[:, 0, 0]) and np.all(P[:, 4, 1] > P[:, 2, 1])


@pytest.mark.parametrize("h", [(0.5, 0.25, 0.25), (1.0 / 8, 1.0 / 4, 1.0 / 3)])
def test_hex27_closed_form_element_matrices(h):
    coords = np.array([[a * h[0] / 2, b * h[1] / 2, c * h[2] / 2] for a, b, c in _POS])
    wq, N, Nd = el.hex27_shape_functions(coords)
    assert np.allclose(N.sum(axis=1), 1.0) and np.allclose(Nd.sum(axis=2), 0.0, atol=1e-12)
    Me = el.mass_element(wq, N)
    Le = el.laplacian_element_calc(wq, Nd)
    Ge = el.gradient_element_calc(wq, N, Nd, 3)
    (Mx, Kx, Gx), (My, Ky, Gy), (Mz, Kz, Gz) = _one_d(h[0]), _one_d(h[1]), _one_d(h[2])
    vol = h[0] * h[1] * h[2]
    assert np.allclose(Me, _tensor(Mx, My, Mz), rtol=0, atol=1e-15 * vol)
    Lcf = _tensor(Kx, My, Mz) + _tensor(Mx, Ky, Mz) + _tensor(Mx, My, Kz)
    assert np.allclose(Le, Lcf, rtol=0, atol=1e-13 * np.abs(Lcf).max())
    Gcf = [_tensor(Gx, My, Mz), _tensor(Mx, Gy, Mz), _tensor(Mx, My, Gz)]
    for d in range(3):
        assert np.allclose(Ge[d], Gcf[d], rtol=0, atol=1e-14 * np.abs(Gcf[d]).max())
    assert np.isclose(Me.sum(), vol)


def test_hex27_convection_matrix_brute_force_quadrature():
    """Ce on hex27 for a velocity linear in x, y, z from an independent 5x5x5 Gauss rule equals
    the oracle's 3x3x3 rule (exact for this integrand, degree <= 5 per direction)."""
    ne3, L = (2, 2, 1), (1.0, 1.0, 0.5)
    coords, nconn, *_ = ms.create_mesh_3d_hex27(*ne3, *L)
    nn = coords.shape[0]
    h = [L[i] / ne3[i] for i in range(3)]
    A = RNG.standard_normal((3, 4))
    U = np.concatenate([A[d, 0] + coords @ A[d, 1:] for d in range(3)])
    wq, N, Nd = el.hex27_shape_functions(coords[nconn[0]])
    x, w = np.polynomial.legendre.leggauss(5)
    x, w = (x + 1) / 2, w / 2
    for beta in (0.5, 1.0):
        Ce = el.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 3, beta)
        for e in (0, 3):
            ul = np.stack([U[nconn[e] + d * nn] for d in range(3)], axis=1)
            ref = np.zeros((27, 27))
            for xa, wa in zip(x, w):
                for xb, wb in zip(x, w):
                    for xc, wc in zip(x, w):
                        Nq = np.array([_B1[a](xa) * _B1[b](xb) * _B1[c](xc) for a, b, c in _POS])
                        dN = np.stack([
                            np.array([_D1[a](xa) * _B1[b](xb) * _B1[c](xc) for a, b, c in _POS]) / h[0],
                            np.array([_B1[a](xa) * _D1[b](xb) * _B1[c](xc) for a, b, c in _POS]) / h[1],
                            np.array([_B1[a](xa) * _B1[b](xb) * _D1[c](xc) for a, b, c in _POS]) / h[2]])
                        uq = Nq @ ul
                        div = sum(dN[d] @ ul[:, d] for d in range(3))
                        jac = wa * wb * wc * h[0] * h[1] * h[2]
                        ref += jac * (np.outer(Nq, uq @ dN) + beta * div * np.outer(Nq, Nq))
            assert np.allclose(Ce[e], ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_hex27_face_integral_simpson_shares():
    """Uniform inflow u = 1 through the x = 0 face (n = -x): each node gets minus its Simpson
    share (1/6, 4/6, 1/6 per direction) of every face element it lies on, and the total is minus
    the face area; y- and z-faces add nothing (the (z, y) mapping of the reference's face
    integral)."""
    n, L = (2, 3, 2), (2.0, 1.5, 1.0)
    s, surf, h, U_d = ns.build(list(n), list(L), False, [0, 1, 2, 3, 4, 5], [], 1e-10, True, "jacobi", None,
                               shape_func_ord=2)
    Ud = ms.set_face_velocity(U_d, surf[3], s.gamma_d, 1.0, s.nn, 1)
    F = el.u_dot_n_boundary_integral_quad9_faces(s.bconn, s.nvec, s.gamma_d, Ud, s.nn, s.coords)
    assert np.isclose(F.sum(), -L[1] * L[2])
    hy, hz = L[1] / n[1], L[2] / n[2]
    simpson = np.array([1.0, 4.0, 1.0]) / 6
    ref = np.zeros(s.nn)
    nnx, nny = 2 * n[0] + 1, 2 * n[1] + 1
    for ke in range(n[2]):
        for je in range(n[1]):
            for c in range(3):
                for b in range(3):
                    node = (2 * ke + c) * nnx * nny + (2 * je + b) * nnx
                    ref[node] -= simpson[b] * simpson[c] * hy * hz
    assert np.allclose(F, ref, atol=1e-14)


def test_hex27_matvecs_match_assembly_and_adjoint():
    s, *_ = ns.build([3, 2, 2], [1.5, 1.0, 1.0], False, [0, 1, 2, 3, 4, 5], [], 1e-10, True, "jacobi", None,
                     shape_func_ord=2)
    nn, nconn = s.nn, s.nconn
    Lc = emult.assemble(nconn, nn, s.Le).tocsr()
    P = RNG.standard_normal(nn)
    U = RNG.standard_normal(3 * nn)
    assert np.allclose(emult.laplacian_element_mult(P, s.Le, nconn, nn), Lc @ P, atol=1e-12)
    GP = emult.gradient_element_mult(P, s.Ge, nconn, nn, 3)
    DU = emult.divergence_element_mult(U, s.Ge, nconn, nn, 3)
    assert np.isclose(GP @ U, P @ DU, rtol=1e-12)
    assert np.isclose(s.M[:nn].sum(), 1.5) and sp.issparse(Lc)
    assert np.allclose(Lc @ np.ones(nn), 0, atol=1e-12)


@pytest.mark.parametrize("precond", ["jacobi", "approx_inv"])
def test_hex27_step_keeps_uniform_flow(precond):
    s, surf, h, U_d = ns.build([3, 2, 2], [1.5, 1.0, 1.0], False, [0, 1, 2, 3, 4, 5], [], 1e-12, True,
                               precond, None, shape_func_ord=2)
    g = s.gamma_d[s.gamma_d < s.nn]
    U_d = ms.set_face_velocity(U_d, g, s.gamma_d, 1.0, s.nn, 1)
    U0 = np.zeros(3 * s.nn)
    U0[:s.nn] = 1.0
    F_p, _ = ns.compute(s, U_d, None, 0, 0)
    U, P, ui, pi = ns.four_step_timestep(s, U0, np.zeros(s.nn), 1e-2, 1.0, 1e-2, U_d, F_p, 0)
    assert np.allclose(U, U0, atol=1e-11) and np.allclose(P, 0, atol=1e-11)
