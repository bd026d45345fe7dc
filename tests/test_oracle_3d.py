"""Pins of the oracle's 3D (hex8) pieces: closed-form element constants, the convection
matrices against an independent higher-order quadrature, the face boundary integral, the
assembled matvecs and the projection step on a flow it must leave unchanged."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import elements as el
from oracle import emult
from oracle import mesh as ms
from oracle import ns

RNG = np.random.default_rng(3003)

# 1D linear-element integrals on [0, h]: mass, stiffness and int N_a dN_b / dx
_M1 = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6
_K1 = np.array([[1.0, -1.0], [-1.0, 1.0]])
_G1 = np.array([[-1.0, 1.0], [-1.0, 1.0]]) / 2
_BX = np.array([0, 1, 1, 0, 0, 1, 1, 0])
_BY = np.array([0, 0, 1, 1, 0, 0, 1, 1])
_BZ = np.array([0, 0, 0, 0, 1, 1, 1, 1])


def _tensor(ax, ay, az):
    """8 x 8 hex matrix from 1D factors: A[a][b] = ax[xa][xb] ay[ya][yb] az[za][zb]."""
    return (ax[np.ix_(_BX, _BX)] * ay[np.ix_(_BY, _BY)] * az[np.ix_(_BZ, _BZ)])


@pytest.mark.parametrize("h", [(1.0 / 8, 1.0 / 4, 1.0 / 4), (2.0 / 128, 1.0 / 64, 1.0 / 64)])
def test_hex8_closed_form_element_matrices(h):
    hx, hy, hz = h
    coords = np.array([[_BX[b] * hx, _BY[b] * hy, _BZ[b] * hz] for b in range(8)])
    wq, N, Nd = el.hex8_shape_functions(coords)
    Me = el.mass_element(wq, N)
    Le = el.laplacian_element_calc(wq, Nd)
    Ge = el.gradient_element_calc(wq, N, Nd, 3)
    M1x, M1y, M1z = _M1 * hx, _M1 * hy, _M1 * hz
    Me_cf = _tensor(M1x, M1y, M1z)
    Le_cf = _tensor(_K1 / hx, M1y, M1z) + _tensor(M1x, _K1 / hy, M1z) + _tensor(M1x, M1y, _K1 / hz)
    G_cf = [_tensor(_G1, M1y, M1z), _tensor(M1x, _G1, M1z), _tensor(M1x, M1y, _G1)]
    scale = hx * hy * hz
    assert np.allclose(Me, Me_cf, rtol=0, atol=1e-15 * scale)
    assert np.allclose(Le, Le_cf, rtol=1e-13, atol=1e-16)
    for d in range(3):
        assert np.allclose(Ge[d], G_cf[d], rtol=1e-13, atol=1e-18)
    assert np.allclose(Le @ np.ones(8), 0, atol=1e-14 * np.abs(Le).max())
    assert np.isclose(Me.sum(), scale)


def test_hex8_convection_matrix_brute_force_quadrature():
    """Ce on hex8 from a 4x4x4 Gauss rule written independently equals the 2x2x2 oracle
    (2-point Gauss is exact for the trilinear integrand in each direction)."""
    ne3 = (3, 2, 2)
    L = (1.5, 1.0, 0.5)
    coords, nconn, *_ = ms.create_mesh_3d(*ne3, *L)
    nn = coords.shape[0]
    h = [L[i] / ne3[i] for i in range(3)]
    U = RNG.standard_normal(3 * nn)
    wq, N, Nd = el.hex8_shape_functions(coords[nconn[0]])
    sx, sy, sz = 2 * _BX - 1.0, 2 * _BY - 1.0, 2 * _BZ - 1.0
    x, w = np.polynomial.legendre.leggauss(4)
    for beta in (0.5, 1.0):
        Ce = el.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 3, beta)
        for e in (0, 5, 11):
            ul = np.stack([U[nconn[e] + d * nn] for d in range(3)], axis=1)
            ref = np.zeros((8, 8))
            for a, wa in zip(x, w):
                for b, wb in zip(x, w):
                    for c, wc in zip(x, w):
                        fx, fy, fz = 1 + sx * a, 1 + sy * b, 1 + sz * c
                        Nq = fx * fy * fz / 8
                        dN = np.stack([sx * fy * fz / 8 * (2 / h[0]), fx * sy * fz / 8 * (2 / h[1]),
                                       fx * fy * sz / 8 * (2 / h[2])])
                        uq = Nq @ ul
                        div = sum(dN[d] @ ul[:, d] for d in range(3))
                        jac = wa * wb * wc * h[0] * h[1] * h[2] / 8
                        ref += jac * (np.outer(Nq, uq @ dN) + beta * div * np.outer(Nq, Nq))
            assert np.allclose(Ce[e], ref, atol=1e-14)


def _channel(n=(4, 2, 2), L=(2.0, 1.0, 1.0), precond="jacobi", tol=1e-10):
    s, surf, h, U_d = ns.build(list(n), list(L), False, [0, 2, 3, 4, 5], [1], tol, True, precond, None)
    return s, surf, U_d


def test_face_integral_flux_and_quirk():
    """Uniform inflow u = 1 on the left face (n = (-1,0,0)): the node values sum to minus the
    face area and equal -(area share) per node.  A velocity normal to a face of constant y
    adds nothing (the reference maps every face from its (z, y) coordinates)."""
    n, L = (4, 3, 2), (2.0, 1.5, 1.0)
    s, surf, U_d = _channel(n, L)
    Ud = ms.set_face_velocity(U_d, surf[3], s.gamma_d, 1.0, s.nn, 1)
    F_p = el.u_dot_n_boundary_integral_2d(s.bconn, s.nvec, s.gamma_d, Ud, s.nn, s.coords)
    assert np.isclose(F_p.sum(), -L[1] * L[2])
    hy, hz = L[1] / n[1], L[2] / n[2]
    y, z = s.coords[surf[3], 1], s.coords[surf[3], 2]
    wy = np.where((y == 0) | np.isclose(y, L[1]), hy / 2, hy)
    wz = np.where((z == 0) | np.isclose(z, L[2]), hz / 2, hz)
    assert np.allclose(F_p[surf[3]], -wy * wz)
    other = np.setdiff1d(np.arange(s.nn), surf[3])
    assert np.allclose(F_p[other], 0.0)
    Ud2 = ms.set_face_velocity(U_d, surf[0], s.gamma_d, 1.0, s.nn, 2)   # v = 1 on the bottom face
    assert np.allclose(el.u_dot_n_boundary_integral_2d(s.bconn, s.nvec, s.gamma_d, Ud2, s.nn, s.coords), 0)


def _assembled(nconn, nn, A):
    rows = np.repeat(nconn, 8, axis=1).ravel()
    cols = np.tile(nconn, (1, 8)).ravel()
    return sp.coo_matrix((np.tile(A.ravel(), nconn.shape[0]), (rows, cols)), shape=(nn, nn)).tocsr()


def test_hex8_matvecs_match_assembly_and_adjoint():
    s, surf, U_d = _channel((3, 2, 2), (1.5, 1.0, 1.0))
    nn, nconn = s.nn, s.nconn
    U, P = RNG.standard_normal(3 * nn), RNG.standard_normal(nn)
    Mc, Lc = _assembled(nconn, nn, s.Me), _assembled(nconn, nn, s.Le)
    got = emult.consistent_mass_mult(U, s.Me, nconn, nn, 3)
    ref = np.concatenate([Mc @ U[d * nn:(d + 1) * nn] for d in range(3)])
    assert np.allclose(got, ref, atol=1e-14)
    assert np.allclose(emult.laplacian_element_mult(P, s.Le, nconn, nn), Lc @ P, atol=1e-13)
    GP = emult.gradient_element_mult(P, s.Ge, nconn, nn, 3)
    DU = emult.divergence_element_mult(U, s.Ge, nconn, nn, 3)
    assert np.isclose(GP @ U, P @ DU, rtol=1e-12)
    M = s.M[:nn]
    assert np.isclose(M.sum(), 1.5 * 1.0 * 1.0)


def test_hex8_step_keeps_uniform_flow():
    """U = (1, 0, 0) everywhere, imposed on every face: convection and viscosity of a uniform
    field vanish and D U is balanced by F_p on the x-faces, so the step returns U unchanged
    and P = 0."""
    s, surf, h, U_d = ns.build([4, 3, 2], [2.0, 1.5, 1.0], False, [0, 1, 2, 3, 4, 5], [], 1e-12, True,
                               "jacobi", None)
    g = s.gamma_d[s.gamma_d < s.nn]
    U_d = ms.set_face_velocity(U_d, g, s.gamma_d, 1.0, s.nn, 1)
    U0 = np.zeros(3 * s.nn)
    U0[:s.nn] = 1.0
    F_p, _ = ns.compute(s, U_d, None, 0, 0)
    U, P, ui, pi = ns.four_step_timestep(s, U0, np.zeros(s.nn), 1e-2, 1.0, 1e-2, U_d, F_p, 0)
    assert np.allclose(U, U0, atol=1e-12) and np.allclose(P, 0, atol=1e-12)


def test_hex8_channel_step_is_sane():
    """A few channel steps from rest with a parabolic inflow: boundary values exact, the
    outlet pressure pinned, iteration counts positive and the flow entering the domain."""
    s, surf, U_d = _channel((8, 4, 4), (2.0, 1.0, 1.0), "jacobi", 1e-10)
    U_d = ns.channel_inflow(s, surf, U_d, 1.0, 1.0)
    U = np.zeros(3 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    for _ in range(3):
        F_p, F = ns.compute(s, U_d, None, 0, 0)
        U, P, ui, pi = ns.four_step_timestep(s, U, P, 2e-3, 1.0, 1e-3, U_d, F_p, F)
        assert ui > 0 and pi > 0
    assert np.all(U[s.gamma_d] == U_d) and np.all(P[surf[1]] == 0)
    assert U[:s.nn].max() > 0.5 and np.isfinite(U).all() and np.isfinite(P).all()


# ----------------------------------------------------------------------------- 3D multigrid
from oracle import multigrid as mg  # noqa: E402


def _p1_dense(nc):
    """1D linear prolongation as a dense (2 nc - 1) x nc matrix, from its definition."""
    P = np.zeros((2 * nc - 1, nc))
    for i in range(nc):
        P[2 * i, i] = 1.0
    for i in range(nc - 1):
        P[2 * i + 1, i] = P[2 * i + 1, i + 1] = 0.5
    return P


def _lv(ne):
    nex, ney, nez = ne
    return dict(ndim=3, nnx=nex + 1, nny=ney + 1, nnz=nez + 1)


def test_trilinear_transfers_are_kron_products_and_adjoint():
    """P3 = Pz (x) Py (x) Px built independently from 1D matrices; R3 = P3^T exactly in
    exact arithmetic (checked to rounding); P3 reproduces trilinear functions."""
    Lc, Lf = _lv((2, 1, 3)), _lv((4, 2, 6))
    nc = Lc["nnx"] * Lc["nny"] * Lc["nnz"]
    nf = Lf["nnx"] * Lf["nny"] * Lf["nnz"]
    P3 = np.kron(_p1_dense(Lc["nnz"]), np.kron(_p1_dense(Lc["nny"]), _p1_dense(Lc["nnx"])))
    assert P3.shape == (nf, nc)
    y = RNG.standard_normal(nc)
    x = RNG.standard_normal(nf)
    assert np.allclose(mg.interpolation_operator_3d(y, Lc, Lf), P3 @ y, rtol=0, atol=1e-14)
    assert np.allclose(mg.restriction_operator_3d(x, Lf, Lc), P3.T @ x, rtol=0, atol=1e-13)
    # trilinear function on the coarse nodes -> exact values at the fine nodes
    k, j, i = np.meshgrid(np.arange(Lc["nnz"]), np.arange(Lc["nny"]), np.arange(Lc["nnx"]), indexing="ij")
    f = lambda i, j, k: 1.5 - 0.5 * i + 2 * j + 0.25 * k + 0.75 * i * j - k * i + 0.125 * i * j * k
    kf, jf, if_ = np.meshgrid(np.arange(Lf["nnz"]), np.arange(Lf["nny"]), np.arange(Lf["nnx"]), indexing="ij")
    fine = mg.interpolation_operator_3d(f(i, j, k).ravel().astype(float), Lc, Lf)
    assert np.allclose(fine, f(if_ / 2, jf / 2, kf / 2).ravel(), rtol=0, atol=1e-13)
    # interior row weight of R: (1 + 2 * 0.5)^3
    assert mg.restriction_operator_3d(np.ones(nf), Lf, Lc).reshape(Lc["nnz"], Lc["nny"], Lc["nnx"])[1, 0, 1] \
        == pytest.approx(8.0 * (1.5 / 2.0))  # y has only 2 coarse nodes: edge weight 1 + 0.5


def test_multigrid_setup_3d_levels():
    s, surf, h, U_d = ns.build([8, 4, 4], [2.0, 1.0, 1.0], False, [0, 2, 3, 4, 5], [1], 1e-10, True, "multigrid",
                               [2, 1, 1])
    assert [(L["nex"], L["ney"], L["nez"]) for L in s.levels] == [(8, 4, 4), (4, 2, 2), (2, 1, 1)]
    for L in s.levels:
        # Le of a box element of size h = (2/nex, 1/ney, 1/nez) (scales with h in 3D)
        hx, hy, hz = 2.0 / L["nex"], 1.0 / L["ney"], 1.0 / L["nez"]
        ref = _tensor(_K1 / hx, _M1 * hy, _M1 * hz) + _tensor(_M1 * hx, _K1 / hy, _M1 * hz) \
            + _tensor(_M1 * hx, _M1 * hy, _K1 / hz)
        assert np.allclose(L["Le"], ref, rtol=0, atol=1e-14)
        # outlet pins: the x = 2 face of the level's mesh
        kk, rr = np.divmod(L["gamma_p"], L["nnx"] * L["nny"])
        assert np.all(rr % L["nnx"] == L["nnx"] - 1) and len(L["gamma_p"]) == L["nny"] * L["nnz"]


def test_two_level_vcycle_3d_dense_restatement():
    """Two-level pinned 3D V-cycle written with dense matrices (Jacobi from zero, residual,
    R = P^T from the 1D kron factors, exact pinned coarse solve, P, Jacobi) equals the oracle's
    V-cycle run with a coarse CG tolerance of 1e-14."""
    s, *_ = ns.build([4, 2, 2], [2.0, 1.0, 1.0], False, [0, 2, 3, 4, 5], [1], 1e-14, True, "multigrid",
                     [2, 1, 1])
    f, c = s.levels
    Af = emult.assemble(f["nconn"], f["nn"], f["Le"]).toarray()
    Ac = emult.assemble(c["nconn"], c["nn"], c["Le"]).toarray()
    P = np.kron(_p1_dense(c["nnz"]), np.kron(_p1_dense(c["nny"]), _p1_dense(c["nnx"])))
    gf, gc = f["gamma_p"], c["gamma_p"]
    Df = np.diag(Af)
    r = RNG.standard_normal(f["nn"])
    r[gf] = 0
    x1 = 2 / 3 * r / Df
    x1[gf] = 0
    Ax = Af @ x1
    Ax[gf] = 0
    rH = P.T @ (r - Ax)
    Acp = Ac.copy()
    Acp[gc, :] = 0
    Acp[:, gc] = 0
    Acp[gc, gc] = 1
    rHp = rH.copy()
    rHp[gc] = 0
    xH = np.linalg.solve(Acp, rHp)
    x = x1 + P @ xH
    x[gf] = 0
    x = 2 / 3 * (r - (Af - np.diag(Df)) @ x) / Df + 1 / 3 * x
    x[gf] = 0
    assert np.allclose(s.M_op(r), x, rtol=0, atol=1e-10 * np.abs(x).max())


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_multigrid_3d_preconditions_pcg(set_zeroP):
    """3D PCG with the V-cycle converges in a handful of iterations where Jacobi needs many,
    and the solution satisfies the pinned / mean-projected system (pinned: outlet face;
    Neumann: closed box)."""
    from oracle.solvers import cg
    outlets = [1] if set_zeroP else []
    res = {}
    for pc in ("multigrid", "jacobi"):
        s, *_ = ns.build([16, 8, 8], [2.0, 1.0, 1.0], False, [0, 2, 3, 4, 5], outlets, 1e-8, set_zeroP, pc,
                         [2, 1, 1])
        A = emult.assemble(s.nconn, s.nn, s.Le).tocsr()
        b = RNG.standard_normal(s.nn) if pc == "multigrid" else res["b"]
        if set_zeroP:
            b[s.gamma_p] = 0
        else:
            b = b - b.mean()

        def op(x):
            y = A @ x
            if set_zeroP:
                y[s.gamma_p] = 0
            return y
        M = s.M_op if set_zeroP else (lambda r: (lambda z: z - z.mean())(s.M_op(r - r.mean())))
        x, k = cg(op, b, tol=1e-8, M=M)
        res[pc] = k
        res["b"] = b
        r = b - op(x)
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(b)
    assert res["multigrid"] < 12 < res["jacobi"]
