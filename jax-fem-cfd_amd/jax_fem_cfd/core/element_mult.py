"""Element-by-element matvecs on the GPU.

Same functions and arguments as jax_fem_cfd/core/element_mult.py; each call is
one (or, per component, a few) launches of the node-owner HIP kernels of
libfcfd (deterministic gather instead of segment_sum).  ``nconn`` must be one
of the reference's structured Q1 meshes -- the kernels use its closed-form
index math, and the grid is recognised from ``nconn`` and ``nn``.
Vectors come back as float64 CUDA tensors.
"""
import numpy as np

from .. import _native as nat
from .._context import grid_context


def _ctx(nconn, nn):
    return grid_context(nconn, int(nn))


def _T(A):
    return np.ascontiguousarray(nat.to_numpy(A).T)


def consistent_mass_mult(S, Me, nconn, nn, ndim):
    """M_c S per component.  element_mult.py:11-36."""
    c = _ctx(nconn, nn)
    S = nat.as_dev(S)
    out = c._empty(ndim * nn)
    for d in range(ndim):
        c.elem_mult_const(Me, S[d * nn:(d + 1) * nn], out=out[d * nn:(d + 1) * nn])
    return out


def laplacian_element_mult(P, Le, nconn, nn):
    """L P.  element_mult.py:38-49."""
    return _ctx(nconn, nn).elem_mult_const(Le, P)


def gradient_element_mult(P, Ge, nconn, nn, ndim):
    """[G_0 P; G_1 P].  element_mult.py:51-77."""
    c = _ctx(nconn, nn)
    P = nat.as_dev(P)
    Ge = nat.to_numpy(Ge)
    out = c._empty(ndim * nn)
    for i in range(ndim):
        c.elem_mult_const(Ge[i], P, out=out[i * nn:(i + 1) * nn])
    return out


def divergence_element_mult(U, Ge, nconn, nn, ndim):
    """sum_i G_i^T U_i.  element_mult.py:133-157."""
    c = _ctx(nconn, nn)
    U = nat.as_dev(U)
    Ge = nat.to_numpy(Ge)
    out = c.elem_mult_const(_T(Ge[0]), U[:nn])
    for i in range(1, ndim):
        c.elem_mult_const(_T(Ge[i]), U[i * nn:(i + 1) * nn], out=out, accumulate=True)
    return out


def visc_conv_mult(U, Ce_all, Le, mu, rho, nconn, nn, ndim):
    """(rho Ce + mu Le) U per component.  element_mult.py:79-104."""
    c = _ctx(nconn, nn)
    return c.elem_mult_stored(U, K0=mu * nat.to_numpy(Le), A1=Ce_all, c1=rho, ncomp=ndim)


def stabilized_visc_conv_mult(U, Ce_all, Se_all, Le, mu, rho, nconn, nn, ndim):
    """(rho Ce + mu Le + Se) U per component.  element_mult.py:106-131."""
    c = _ctx(nconn, nn)
    return c.elem_mult_stored(U, K0=mu * nat.to_numpy(Le), A1=Ce_all, c1=rho, A2=Se_all, c2=1.0, ncomp=ndim)


def scalar_visc_conv_mult(phi, Ce_all, Le, D, nconn, nn):
    """(Ce + D Le) phi.  element_mult.py:159-172."""
    return _ctx(nconn, nn).elem_mult_stored(phi, K0=D * nat.to_numpy(Le), A1=Ce_all, c1=1.0)


def scalar_stabilized_visc_conv_mult(phi, Ce_all, Se_all, Le, D, nconn, nn):
    """(Ce + D Le + Se) phi.  element_mult.py:174-187."""
    return _ctx(nconn, nn).elem_mult_stored(phi, K0=D * nat.to_numpy(Le), A1=Ce_all, c1=1.0, A2=Se_all, c2=1.0)


def scalar_consistent_mass_SUPG_mult(phi, Me_tau_all, Me, nconn, nn):
    """(Me + Me_tau) phi.  element_mult.py:190-201."""
    return _ctx(nconn, nn).elem_mult_stored(phi, K0=nat.to_numpy(Me), A1=Me_tau_all, c1=1.0)
