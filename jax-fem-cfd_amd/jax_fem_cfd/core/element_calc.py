"""Element matrices.

Same functions as jax_fem_cfd/core/element_calc.py.  The shared 4x4 constants
(Me, Le, Ge) are set-up data computed on the host from the Gauss data; the
per-node sums (lumped mass, L_diag), the per-element convection matrices and
the boundary integral run on the GPU.
"""
import numpy as np

from .. import _native as nat
from .._context import Context, _h_from_shape, grid_context


def _mass_elem(wq, N):
    return np.einsum("q,qa,qb->ab", wq[:, 0], N, N)


def lumped_mass_matrix(nconn, wq, N, nn, ndim):
    """element_calc.py:9-25 -> (M (ndim*nn,) on the GPU, Me)."""
    wq, N = nat.to_numpy(wq), nat.to_numpy(N)
    Me = _mass_elem(wq, N)
    c = grid_context(nconn, int(nn))
    ones = nat.as_dev(np.ones(int(nn)))
    M = c.elem_mult_const(Me, ones)
    return nat.torch().cat([M] * ndim), Me


def laplacian_element_calc(wq, Nderiv):
    """element_calc.py:27-32: Le = sum_q w dN^T dN."""
    wq, Nderiv = nat.to_numpy(wq), nat.to_numpy(Nderiv)
    return np.einsum("q,qda,qdb->ab", wq[:, 0], Nderiv, Nderiv)


def gradient_element_calc(wq, N, Nderiv, ndim):
    """element_calc.py:34-40: Ge[i] = sum_q w N (d_i N)^T."""
    wq, N, Nderiv = nat.to_numpy(wq), nat.to_numpy(N), nat.to_numpy(Nderiv)
    return np.stack([np.einsum("q,qa,qb->ab", wq[:, 0], N, Nderiv[:, i, :]) for i in range(ndim)])


def laplacian_diag(Le, nconn, nn):
    """element_calc.py:99-109: scatter of diag(Le) -> (nn,) on the GPU."""
    c = grid_context(nconn, int(nn))
    D = np.diag(np.diag(nat.to_numpy(Le)))
    return c.elem_mult_const(D, nat.as_dev(np.ones(int(nn))))


def conservative_convec_elem_calc(nconn, U, wq, N, Nderiv, nn, ndim, beta):
    """element_calc.py:165-185: Ce = Ce_A + beta Ce_D per element, (ne, 4, 4)
    float64 CUDA tensor (one HIP thread per element)."""
    if ndim != 2:
        raise NotImplementedError("per-element convection matrices are built for 2D Q1 only")
    hx, hy = _h_from_shape(wq, Nderiv)
    c = grid_context(nconn, int(nn), hx, hy)
    return c.conv_elem_calc(U, beta)


_bi_cache = {}


def u_dot_n_boundary_integral_1d(bconn, nvec, gamma_d, U_d, nn, node_coords, shapefunc=None):
    """element_calc.py:111-135: F_p = int q (u.n) over the Dirichlet edges."""
    coords = nat.to_numpy(node_coords)
    nn = int(nn)
    xs = np.unique(coords[:, 0])
    ys = np.unique(coords[:, 1])
    periodic = False
    nex, ney = len(xs) - 1, len(ys) - 1
    key = (nex, ney, float(xs[-1]), float(ys[-1]), nat.to_numpy(bconn).tobytes(), nat.to_numpy(gamma_d).tobytes())
    ctx = _bi_cache.get(key)
    if ctx is None:
        ctx = Context(nex, ney, float(xs[-1]), float(ys[-1]), periodic, gamma_d=nat.to_numpy(gamma_d),
                      bconn=nat.to_numpy(bconn), nvec=nat.to_numpy(nvec))
        _bi_cache[key] = ctx
    if ctx.nn != nn:
        raise ValueError("node_coords do not describe the structured grid of nn nodes")
    return ctx.boundary_integral(U_d)


def Ctau_tilde_elem_calc(nconn, U, wq, N, Nderiv, nn, ndim):
    """element_calc.py:188-200: Ct[a,b] = sum_q w (U_q . grad N_a)(U_q . grad N_b), (ne, 4, 4)."""
    if ndim != 2:
        raise NotImplementedError("SUPG element matrices are built for 2D Q1 only")
    hx, hy = _h_from_shape(wq, Nderiv)
    Ct, _ = grid_context(nconn, int(nn), hx, hy).supg_elem_calc(U)
    return Ct


def calculate_tau(nconn, U, Ce_all, Ctau_tilde_all, N, nn, ndim, D, dt, r):
    """element_calc.py:202-229: SUPG tau per element from |Ce|, |Ct| (the latter recomputed
    from U on the device) and the quadrature velocities of U."""
    c = grid_context(nconn, int(nn))
    _, tau = c.supg_elem_calc(U, Ce=Ce_all, D=D, dt=dt, r=r)
    return tau
