"""Element-by-element matvecs: gather -> dense element product -> scatter-add
(oracle, test-only).  Follows jax_fem_cfd/core/element_mult.py.

The scatter is ``np.bincount`` over ``nconn.ravel()`` which accumulates in
element order, the same order as a sequential segment_sum.
"""
import numpy as np


def _scatter(data, nconn, nn):
    return np.bincount(nconn.ravel(), data.ravel(), minlength=nn)


def _const_mult(x, A, nconn, nn):
    return _scatter(x[nconn] @ A.T, nconn, nn)


def consistent_mass_mult(S, Me, nconn, nn, ndim):
    """M_c S per component.  element_mult.py:11-36 (the ndim=3 branch of the
    reference raises NameError on an undefined y_result; here it is the
    evident per-component product)."""
    return np.concatenate([_const_mult(S[d * nn:(d + 1) * nn], Me, nconn, nn) for d in range(ndim)])


def laplacian_element_mult(P, Le, nconn, nn):
    """L P.  element_mult.py:38-49 (also used with Le_off by the smoother)."""
    return _const_mult(P, Le, nconn, nn)


def gradient_element_mult(P, Ge, nconn, nn, ndim):
    """G P = [G_0 P; G_1 P(; G_2 P)].  element_mult.py:51-77."""
    return np.concatenate([_const_mult(P, Ge[i], nconn, nn) for i in range(ndim)])


def divergence_element_mult(U, Ge, nconn, nn, ndim):
    """D U = sum_i G_i^T U_i.  element_mult.py:133-157."""
    out = _const_mult(U[:nn], Ge[0].T, nconn, nn)
    for i in range(1, ndim):
        out = out + _const_mult(U[i * nn:(i + 1) * nn], Ge[i].T, nconn, nn)
    return out


def visc_conv_mult(U, Ce_all, Le, mu, rho, nconn, nn, ndim):
    """(rho Ce + mu Le) U per component.  element_mult.py:79-104."""
    Ke = mu * Le
    A = rho * Ce_all + Ke[None]
    outs = []
    for d in range(ndim):
        ue = U[nconn + d * nn]
        outs.append(_scatter(np.einsum("eab,eb->ea", A, ue), nconn, nn))
    return np.concatenate(outs)


def scalar_visc_conv_mult(phi, Ce_all, Le, D, nconn, nn):
    """(Ce + D Le) phi.  element_mult.py:159-172."""
    A = Ce_all + (D * Le)[None]
    return _scatter(np.einsum("eab,eb->ea", A, phi[nconn]), nconn, nn)


def scalar_stabilized_visc_conv_mult(phi, Ce_all, Se_all, Le, D, nconn, nn):
    """(Ce + D Le + Se) phi.  element_mult.py:174-187."""
    A = Ce_all + (D * Le)[None] + Se_all
    return _scatter(np.einsum("eab,eb->ea", A, phi[nconn]), nconn, nn)


def scalar_consistent_mass_SUPG_mult(phi, Me_tau_all, Me, nconn, nn):
    """(Me + Me_tau) phi.  element_mult.py:190-201."""
    A = Me[None] + Me_tau_all
    return _scatter(np.einsum("eab,eb->ea", A, phi[nconn]), nconn, nn)


def assemble(nconn, nn, elem_mats):
    """scipy.sparse CSR assembly of per-element (ne, I, I) or shared (I, I)
    matrices -- the independent check for every matvec above."""
    import scipy.sparse as sp
    ne, nen = nconn.shape
    if elem_mats.ndim == 2:
        elem_mats = np.broadcast_to(elem_mats, (ne, nen, nen))
    rows = np.repeat(nconn, nen, axis=1).ravel()
    cols = np.tile(nconn, (1, nen)).ravel()
    return sp.coo_matrix((elem_mats.reshape(ne, -1).ravel(), (rows, cols)), shape=(nn, nn)).tocsr()
