"""Shape functions and element matrices (oracle, test-only).

Follows jax_fem_cfd/core/shape_functions.py and jax_fem_cfd/core/element_calc.py.
Index convention as in the reference: I = element node, q = quadrature
point, d = spatial dimension.  N is (q, I), Nderiv is (q, d, I), wq is (q, 1).
"""
import numpy as np

_S3 = 1.0 / np.sqrt(3.0)


def lin2_shape_functions(coords):
    """2-node line, 2-point Gauss.  shape_functions.py:11-25."""
    xi = np.array([-1.0, 1.0]) * _S3
    N = np.stack([(1 - xi) / 2, (1 + xi) / 2], axis=1)
    dNdxi = np.array([[-0.5, 0.5], [-0.5, 0.5]])
    d = coords[1] - coords[0]
    h = np.sqrt(d[0] ** 2 + d[1] ** 2)
    Nderiv = (dNdxi * (2 / h))[:, None, :]
    wq = (np.ones(2) * (h / 2))[:, None]
    return wq, N, Nderiv


def quad4_shape_functions(coords):
    """Bilinear quad, 2x2 Gauss, q-order (-,-),(+,-),(+,+),(-,+); wq = det J.
    shape_functions.py:27-45."""
    xi = np.array([-1.0, 1.0, 1.0, -1.0]) * _S3
    eta = np.array([-1.0, -1.0, 1.0, 1.0]) * _S3
    N = np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta),
                  (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)], axis=1) / 4
    dNdxi = np.stack([-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)], axis=1) / 4
    dNdeta = np.stack([-(1 - xi), -(1 + xi), 1 + xi, 1 - xi], axis=1) / 4
    J = np.stack([dNdxi @ coords, dNdeta @ coords], axis=1)          # (q, 2, 2)
    wq = np.linalg.det(J)[:, None]
    Nderiv = np.linalg.inv(J) @ np.stack([dNdxi, dNdeta], axis=1)     # (q, 2, I)
    return wq, N, Nderiv


def hex8_shape_functions(coords):
    """Trilinear hex, 2x2x2 Gauss.  shape_functions.py:47-86."""
    xi = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float) * _S3
    eta = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float) * _S3
    zeta = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float) * _S3
    sx = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float)
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float)
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float)
    fx = 1 + xi[:, None] * sx[None, :]
    fy = 1 + eta[:, None] * sy[None, :]
    fz = 1 + zeta[:, None] * sz[None, :]
    N = fx * fy * fz / 8
    dNdxi = sx[None, :] * fy * fz / 8
    dNdeta = fx * sy[None, :] * fz / 8
    dNdzeta = fx * fy * sz[None, :] / 8
    J = np.stack([dNdxi @ coords, dNdeta @ coords, dNdzeta @ coords], axis=1)
    wq = np.linalg.det(J)[:, None]
    Nderiv = np.linalg.inv(J) @ np.stack([dNdxi, dNdeta, dNdzeta], axis=1)
    return wq, N, Nderiv


_S35 = np.sqrt(3.0 / 5.0)


def lin3_shape_functions(coords):
    """3-node line, 3-point Gauss (-sqrt(3/5), 0, sqrt(3/5)), weights (5, 8, 5)/9 scaled by
    |dx/dxi|.  shape_functions.py:89-106."""
    xi = np.array([-_S35, 0.0, _S35])
    w = np.array([5.0, 8.0, 5.0]) / 9
    N = np.stack([xi * (xi - 1) / 2, 1 - xi ** 2, xi * (xi + 1) / 2], axis=1)
    dN = np.stack([(2 * xi - 1) / 2, -2 * xi, (2 * xi + 1) / 2], axis=1)
    J = np.sqrt((dN @ coords[:, 0]) ** 2 + (dN @ coords[:, 1]) ** 2)
    return (w * J)[:, None], N, (dN / J[:, None])[:, None, :]


def quad9_shape_functions(coords):
    """Biquadratic 9-node quad, 3x3 Gauss; q = (eta index) * 3 + (xi index), node order of
    create_mesh_2d_quad9.  shape_functions.py:108-159."""
    x1 = np.array([-_S35, 0.0, _S35])
    w1 = np.array([5.0, 8.0, 5.0]) / 9
    xi, eta = np.meshgrid(x1, x1)
    xi, eta = xi.ravel(), eta.ravel()
    wq = np.outer(w1, w1).ravel()
    # 1D quadratic Lagrange factors at local positions 0, 1, 2 (xi = -1, 0, 1) and derivatives
    L = lambda t: [t * (t - 1) / 2, 1 - t ** 2, t * (t + 1) / 2]
    dL = lambda t: [(2 * t - 1) / 2, -2 * t, (2 * t + 1) / 2]
    pos = [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (0, 1), (1, 1)]
    Lx, Ly, dLx, dLy = L(xi), L(eta), dL(xi), dL(eta)
    N = np.stack([Lx[a] * Ly[b] for a, b in pos], axis=1)
    dNdxi = np.stack([dLx[a] * Ly[b] for a, b in pos], axis=1)
    dNdeta = np.stack([Lx[a] * dLy[b] for a, b in pos], axis=1)
    J = np.stack([dNdxi @ coords, dNdeta @ coords], axis=1)
    w = wq * np.linalg.det(J)
    Nderiv = np.linalg.inv(J) @ np.stack([dNdxi, dNdeta], axis=1)
    return w[:, None], N, Nderiv


def hex27_shape_functions(coords):
    """Triquadratic 27-node hex, 3x3x3 Gauss; q = (zeta index * 3 + eta index) * 3 + xi index,
    node b = (c*3 + b_y)*3 + a at local position (a, b_y, c), as create_mesh_3d_hex27.  The
    reference has no 3D quadratic element (simulation.py:73-74); this is the tensor product of
    the 1D quadratic Lagrange factors that quad9_shape_functions (shape_functions.py:108-159)
    uses in 2D."""
    x1 = np.array([-_S35, 0.0, _S35])
    w1 = np.array([5.0, 8.0, 5.0]) / 9
    zeta, eta, xi = np.meshgrid(x1, x1, x1, indexing="ij")
    xi, eta, zeta = xi.ravel(), eta.ravel(), zeta.ravel()
    wz, wy, wx = np.meshgrid(w1, w1, w1, indexing="ij")
    w = (wx * wy * wz).ravel()
    L = lambda t: [t * (t - 1) / 2, 1 - t ** 2, t * (t + 1) / 2]
    dL = lambda t: [(2 * t - 1) / 2, -2 * t, (2 * t + 1) / 2]
    Lx, Ly, Lz, dLx, dLy, dLz = L(xi), L(eta), L(zeta), dL(xi), dL(eta), dL(zeta)
    pos = [(a, b, c) for c in range(3) for b in range(3) for a in range(3)]
    N = np.stack([Lx[a] * Ly[b] * Lz[c] for a, b, c in pos], axis=1)
    dNdxi = np.stack([dLx[a] * Ly[b] * Lz[c] for a, b, c in pos], axis=1)
    dNdeta = np.stack([Lx[a] * dLy[b] * Lz[c] for a, b, c in pos], axis=1)
    dNdzeta = np.stack([Lx[a] * Ly[b] * dLz[c] for a, b, c in pos], axis=1)
    J = np.stack([dNdxi @ coords, dNdeta @ coords, dNdzeta @ coords], axis=1)
    wq = (w * np.linalg.det(J))[:, None]
    Nderiv = np.linalg.inv(J) @ np.stack([dNdxi, dNdeta, dNdzeta], axis=1)
    return wq, N, Nderiv


def quad9_face_shape_functions(coords):
    """Marker for the quad9-faced boundary integral of hex27 meshes (see
    u_dot_n_boundary_integral_quad9_faces); evaluating it is the quad9 shape function."""
    return quad9_shape_functions(coords)


def shape_functions_for(ndim, shape_func_ord=1):
    if ndim == 2 and shape_func_ord == 2:
        return quad9_shape_functions, lin3_shape_functions
    if ndim == 2:
        return quad4_shape_functions, lin2_shape_functions
    if shape_func_ord == 2:
        return hex27_shape_functions, quad9_face_shape_functions
    return hex8_shape_functions, quad4_shape_functions


def mass_element(wq, N):
    """Me[a,b] = sum_q w N_qa N_qb.  element_calc.py:13."""
    return np.einsum("q,qa,qb->ab", wq[:, 0], N, N)


def lumped_mass_matrix(nconn, wq, N, nn, ndim):
    """Row-sum lumped mass scattered to nodes, tiled per component.
    element_calc.py:9-25.  Returns (M (ndim*nn,), Me)."""
    Me = mass_element(wq, N)
    Ml = Me.sum(axis=1)
    M = np.bincount(nconn.ravel(), np.tile(Ml, nconn.shape[0]), minlength=nn)
    return np.concatenate([M] * ndim), Me


def laplacian_element_calc(wq, Nderiv):
    """Le = sum_q w dN^T dN.  element_calc.py:27-32."""
    return np.einsum("q,qda,qdb->ab", wq[:, 0], Nderiv, Nderiv)


def gradient_element_calc(wq, N, Nderiv, ndim):
    """Ge[i][a,b] = sum_q w N_qa d_i N_qb.  element_calc.py:34-40."""
    return np.stack([np.einsum("q,qa,qb->ab", wq[:, 0], N, Nderiv[:, i, :]) for i in range(ndim)])


def laplacian_diag(Le, nconn, nn):
    """Scatter of diag(Le).  element_calc.py:99-109."""
    return np.bincount(nconn.ravel(), np.tile(np.diag(Le), nconn.shape[0]), minlength=nn)


def element_velocity(nconn, U, nn, ndim):
    """U_local (ne, I, ndim) = U[nconn + nn*d]."""
    return np.stack([U[nconn + nn * d] for d in range(ndim)], axis=-1)


def conservative_convec_elem_calc(nconn, U, wq, N, Nderiv, nn, ndim, beta):
    """Per-element convection matrices Ce = Ce_A + beta*Ce_D (Gresho-Sani).
    element_calc.py:165-185.

    Ce_A[a,b] = sum_q w N_qa (U_q . grad N_qb),  U_q = N U_local
    Ce_D[a,b] = sum_q w divU_q N_qa N_qb,        divU_q = sum_dI dN_qdI U_Id
    Returns (ne, I, I).
    """
    Ul = element_velocity(nconn, U, nn, ndim)                 # (e, I, d)
    Uq = np.einsum("qc,ecd->eqd", N, Ul)                      # (e, q, d)
    adv = np.einsum("eqd,qdb->eqb", Uq, Nderiv)               # (e, q, b)
    CeA = np.einsum("q,qa,eqb->eab", wq[:, 0], N, adv)
    divU = np.einsum("qdc,ecd->eq", Nderiv, Ul)               # (e, q)
    CeD = np.einsum("q,eq,qa,qb->eab", wq[:, 0], divU, N, N)
    return CeA + beta * CeD


def u_dot_n_boundary_integral_1d(bconn, nvec, gamma_d, U_d, nn, coords, shapefunc=lin2_shape_functions):
    """F_p = int_Gamma q (u.n), lin2 Gauss per boundary edge, then scatter.
    element_calc.py:111-135."""
    ui = np.searchsorted(gamma_d, bconn)
    vi = np.searchsorted(gamma_d, bconn + nn)
    Ub = np.stack([U_d[ui], U_d[vi]], axis=-1)                # (ne_b, I, 2)
    Fe = np.empty(bconn.shape, dtype=float)
    for e in range(bconn.shape[0]):
        wq, N, _ = shapefunc(coords[bconn[e]])
        Uq = N @ Ub[e]
        un = (Uq * nvec[e][None, :]).sum(axis=1)
        Fe[e] = (N * wq).T @ un
    return np.bincount(bconn.ravel(), Fe.ravel(), minlength=nn)


def u_dot_n_boundary_integral_2d(bconn, nvec, gamma_d, U_d, nn, coords):
    """F_p = int_Gamma q (u.n) over Dirichlet faces of a hex mesh, quad4 2x2 Gauss per face,
    then scatter.  element_calc.py:137-163.  As in the reference every face is mapped from the
    (z, y) coordinates of its nodes (columns [2, 1]), so faces of constant y or z have det J = 0
    and contribute nothing; only w = det J and N are used (J^-1 is not formed, the reference
    discards it)."""
    ui = np.searchsorted(gamma_d, bconn)
    vi = np.searchsorted(gamma_d, bconn + nn)
    wi = np.searchsorted(gamma_d, bconn + 2 * nn)
    Ub = np.stack([U_d[ui], U_d[vi], U_d[wi]], axis=-1)              # (nf, I, 3)
    xi = np.array([-1.0, 1.0, 1.0, -1.0]) * _S3
    eta = np.array([-1.0, -1.0, 1.0, 1.0]) * _S3
    N = np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta),
                  (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)], axis=1) / 4
    dNdxi = np.stack([-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)], axis=1) / 4
    dNdeta = np.stack([-(1 - xi), -(1 + xi), 1 + xi, 1 - xi], axis=1) / 4
    P = coords[bconn][:, :, [2, 1]]                                  # (nf, I, 2)
    Jx = np.einsum("qi,fic->fqc", dNdxi, P)
    Je = np.einsum("qi,fic->fqc", dNdeta, P)
    wq = Jx[..., 0] * Je[..., 1] - Jx[..., 1] * Je[..., 0]           # (nf, q) = det J
    Uq = np.einsum("qi,fid->fqd", N, Ub)
    un = np.einsum("fqd,fd->fq", Uq, nvec.astype(float))
    Fe = np.einsum("qi,fq,fq->fi", N, wq, un)
    return np.bincount(bconn.ravel(), Fe.ravel(), minlength=nn)


def u_dot_n_boundary_integral_quad9_faces(bconn, nvec, gamma_d, U_d, nn, coords):
    """u_dot_n_boundary_integral_2d (element_calc.py:137-163) on the quad9 faces of a hex27
    mesh, 3x3 Gauss per face.  As the reference's 3D face integral does, every face is mapped from
    the (z, y) coordinates of its nodes, so only faces of constant x contribute."""
    ui = np.searchsorted(gamma_d, bconn)
    vi = np.searchsorted(gamma_d, bconn + nn)
    wi = np.searchsorted(gamma_d, bconn + 2 * nn)
    Ub = np.stack([U_d[ui], U_d[vi], U_d[wi]], axis=-1)              # (nf, 9, 3)
    x1 = np.array([-_S35, 0.0, _S35])
    w1 = np.array([5.0, 8.0, 5.0]) / 9
    xi, eta = np.meshgrid(x1, x1)
    xi, eta = xi.ravel(), eta.ravel()
    w = np.outer(w1, w1).ravel()
    L = lambda t: [t * (t - 1) / 2, 1 - t ** 2, t * (t + 1) / 2]
    dL = lambda t: [(2 * t - 1) / 2, -2 * t, (2 * t + 1) / 2]
    pos = [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (0, 1), (1, 1)]
    Lx, Ly, dLx, dLy = L(xi), L(eta), dL(xi), dL(eta)
    N = np.stack([Lx[a] * Ly[b] for a, b in pos], axis=1)
    dNdxi = np.stack([dLx[a] * Ly[b] for a, b in pos], axis=1)
    dNdeta = np.stack([Lx[a] * dLy[b] for a, b in pos], axis=1)
    P = coords[bconn][:, :, [2, 1]]
    Jx = np.einsum("qi,fic->fqc", dNdxi, P)
    Je = np.einsum("qi,fic->fqc", dNdeta, P)
    wq = w[None, :] * (Jx[..., 0] * Je[..., 1] - Jx[..., 1] * Je[..., 0])
    Uq = np.einsum("qi,fid->fqd", N, Ub)
    un = np.einsum("fqd,fd->fq", Uq, nvec.astype(float))
    Fe = np.einsum("qi,fq,fq->fi", N, wq, un)
    return np.bincount(bconn.ravel(), Fe.ravel(), minlength=nn)


def Ctau_tilde_elem_calc(nconn, U, wq, N, Nderiv, nn, ndim):
    """Ct[a,b] = sum_q w dN_qa^T (U_q U_q^T) dN_qb.  element_calc.py:188-200."""
    Ul = element_velocity(nconn, U, nn, ndim)
    Uq = np.einsum("qc,ecd->eqd", N, Ul)
    g = np.einsum("eqd,qda->eqa", Uq, Nderiv)                 # (U_q . grad N_a)
    return np.einsum("q,eqa,eqb->eab", wq[:, 0], g, g)


def calculate_tau(nconn, U, Ce_all, Ctau_tilde_all, N, nn, ndim, D, dt, r):
    """SUPG tau, r-norm blend of (|Ce|/|Ct|, dt/2, Re*|Ce|/|Ct|).
    element_calc.py:202-229."""
    D = D + 1e-6
    Ul = element_velocity(nconn, U, nn, ndim)
    Uq = np.einsum("qc,ecd->eqd", N, Ul)
    ce = np.sqrt((Ce_all ** 2).sum(axis=(1, 2)))
    ct = np.sqrt((Ctau_tilde_all ** 2).sum(axis=(1, 2)))
    u2 = (Uq ** 2).sum(axis=2).mean(axis=1)
    Re = (u2 * ce) / (D * ct)
    t1 = ce / ct
    t2 = dt / 2
    t3 = t1 * Re
    return (1 / t1 ** r + 1 / t2 ** r + 1 / t3 ** r) ** (-1 / r)
