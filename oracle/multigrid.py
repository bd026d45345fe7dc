"""Geometric multigrid V-cycle (oracle, test-only).

Follows jax_fem_cfd/core/multigrid.py.  Levels are kept as a Python list of
dicts instead of the reference's packed arrays with cumulative offsets; the
numbers are the same.
"""
import numpy as np

from . import elements as el
from . import emult
from .solvers import cg


def multigrid_setup(num_elem, num_elemf, domain_size, outlet_faces, mesh_func, shapefunc,
                    partial_dirichlet=None, shape_func_ord=1):
    """Per level: mesh, Le, L_diag, Le_off = Le - diag(Le), gamma_p (outlet
    nodes or [0]).  multigrid.py:9-88.  Level count uses the x ratio only
    (:12) and both directions halve in lockstep (:80).  Three element counts
    give the hex8 hierarchy (see multigrid_setup_3d)."""
    if len(num_elem) == 3:
        return multigrid_setup_3d(num_elem, num_elemf, domain_size, outlet_faces, mesh_func, shapefunc,
                                  shape_func_ord)
    nex, ney = num_elem
    nlev = int(np.log2(nex // num_elemf[0])) + 1
    levels = []
    for _ in range(nlev):
        coords, nconn, surf, bconns, _ = mesh_func(nex, ney, *domain_size)
        wq, _, Nderiv = shapefunc(coords[nconn[0]])
        nn = coords.shape[0]
        outlet = []
        for f in outlet_faces or []:
            outlet.append(surf[f])
        if partial_dirichlet:
            for f, (s_frac, e_frac, rem) in partial_dirichlet.items():
                nfe = bconns[f].shape[0]
                s_idx, e_idx = round(s_frac * nfe), round(e_frac * nfe)
                if shape_func_ord == 2:
                    s_idx, e_idx = 2 * s_idx, 2 * e_idx
                if rem == "outlet":
                    outlet.append(surf[f][:s_idx])
                    outlet.append(surf[f][e_idx + 1:])
        gamma_p = np.unique(np.concatenate(outlet)).astype(np.int64) if outlet else np.array([0])
        Le = el.laplacian_element_calc(wq, Nderiv)
        k = shape_func_ord
        levels.append(dict(nex=nex, ney=ney, nnx=k * nex + 1, nny=k * ney + 1, nn=nn, nconn=nconn,
                           Le=Le, Le_off=Le - np.diag(np.diag(Le)),
                           L_diag=el.laplacian_diag(Le, nconn, nn), gamma_p=gamma_p))
        nex, ney = nex // 2, ney // 2
    return levels


def _pad_periodic(x, nny, nnx):
    """(nny-1, nnx-1) periodic grid -> (nny, nnx) with the wrap row/col."""
    x = x.reshape(nny - 1, nnx - 1)
    out = np.zeros((nny, nnx))
    out[:-1, :-1] = x
    out[:-1, -1] = x[:, 0]
    out[-1, :-1] = x[0, :]
    out[-1, -1] = x[0, 0]
    return out


def interpolation_operator(xH, nnHx, nnHy, nnhx, nnhy, periodic):
    """Bilinear prolongation on node grids (injection, 1/2 and 1/4
    averages).  multigrid.py:90-140; periodic grids are padded with the wrap
    and cropped afterwards."""
    H = _pad_periodic(xH, nnHy, nnHx) if periodic else xH.reshape(nnHy, nnHx)
    h = np.zeros((nnhy, nnhx))
    h[0::2, 0::2] = H
    h[0::2, 1::2] = 0.5 * H[:, :-1] + 0.5 * H[:, 1:]
    h[1::2, 0::2] = 0.5 * H[:-1, :] + 0.5 * H[1:, :]
    h[1::2, 1::2] = 0.25 * H[:-1, :-1] + 0.25 * H[:-1, 1:] + 0.25 * H[1:, :-1] + 0.25 * H[1:, 1:]
    if periodic:
        h = h[:-1, :-1]
    return h.ravel()


def restriction_operator(xh, nnhx, nnhy, nnHx, nnHy, periodic):
    """R = I^T without the 1/4 full-weighting factor, with the explicit
    edge/corner formulas of multigrid.py:142-212.  On periodic grids the
    padded fine grid is restricted with the non-periodic formulas and the
    wrap row/column of the coarse result is dropped (:207-208), so there the
    operator is not the exact transpose of the periodic prolongation."""
    h = _pad_periodic(xh, nnhy, nnhx) if periodic else xh.reshape(nnhy, nnhx)
    hp = np.zeros((nnhy + 2, nnhx + 2))
    hp[1:-1, 1:-1] = h
    c = hp[1:-1:2, 1:-1:2]
    e = hp[1:-1:2, 2::2] + hp[1:-1:2, 0:-2:2] + hp[2::2, 1:-1:2] + hp[0:-2:2, 1:-1:2]
    d = hp[2::2, 2::2] + hp[2::2, 0:-2:2] + hp[0:-2:2, 2::2] + hp[0:-2:2, 0:-2:2]
    H = c + 0.5 * e + 0.25 * d
    if periodic:
        H = H[:-1, :-1]
    return H.ravel()


def single_jacobi_iter(x, rhs, Le_off, L_diag, gamma_p, nconn, nn, set_zeroP):
    """Damped Jacobi, omega = 2/3, then pin or mean-project.
    multigrid.py:214-228."""
    omega = 2.0 / 3.0
    x = x.copy()
    if set_zeroP:
        x[gamma_p] = 0
    off = emult.laplacian_element_mult(x, Le_off, nconn, nn)
    x = (omega * (rhs - off) / L_diag) + (1 - omega) * x
    if set_zeroP:
        x[gamma_p] = 0
    else:
        x = x - x.mean()
    return x


# --------------------------------------------------------------------------------------------
# 3D (hex8) hierarchy.  The reference stops at 2D (simulation.py:75-76 rejects 3D multigrid); this
# is the same V-cycle on box grids: the x ratio sets the level count, all three directions halve
# in lockstep, Le is recomputed from each level's element (it scales with h in 3D), the transfers
# are trilinear prolongation and its exact transpose (separable: x, then y, then z).
# --------------------------------------------------------------------------------------------

def multigrid_setup_3d(num_elem, num_elemf, domain_size, outlet_faces, mesh_func, shapefunc, shape_func_ord=1):
    """Box-grid levels; hex27 levels (shape_func_ord 2) keep their node grids (2 ne + 1 per
    direction), which the node-grid transfers below treat as hex8 grids of 2 ne elements."""
    k = shape_func_ord
    nex, ney, nez = num_elem
    nlev = int(np.log2(nex // num_elemf[0])) + 1
    levels = []
    for _ in range(nlev):
        coords, nconn, surf, _, _ = mesh_func(nex, ney, nez, *domain_size)
        wq, _, Nderiv = shapefunc(coords[nconn[0]])
        nn = coords.shape[0]
        outlet = [surf[f] for f in (outlet_faces or [])]
        gamma_p = np.unique(np.concatenate(outlet)).astype(np.int64) if outlet else np.array([0])
        Le = el.laplacian_element_calc(wq, Nderiv)
        levels.append(dict(ndim=3, nex=nex, ney=ney, nez=nez, nnx=k * nex + 1, nny=k * ney + 1, nnz=k * nez + 1, nn=nn,
                           nconn=nconn, Le=Le, Le_off=Le - np.diag(np.diag(Le)),
                           L_diag=el.laplacian_diag(Le, nconn, nn), gamma_p=gamma_p))
        nex, ney, nez = nex // 2, ney // 2, nez // 2
    return levels


def _prolong_axis(H, axis):
    """Linear interpolation along one axis of a node array: 2n - 1 fine nodes from n coarse ones
    (injection at even fine nodes, 0.5 a + 0.5 b at odd ones)."""
    n = H.shape[axis]
    shape = list(H.shape)
    shape[axis] = 2 * n - 1
    h = np.zeros(shape)
    idx = lambda sl: tuple(sl if a == axis else slice(None) for a in range(H.ndim))
    h[idx(slice(0, None, 2))] = H
    h[idx(slice(1, None, 2))] = 0.5 * H[idx(slice(0, -1))] + 0.5 * H[idx(slice(1, None))]
    return h


def _restrict_axis(h, axis):
    """Transpose of _prolong_axis: c[i] = f[2i] + 0.5 (f[2i-1] + f[2i+1]), missing ends 0."""
    n = (h.shape[axis] + 1) // 2
    idx = lambda sl: tuple(sl if a == axis else slice(None) for a in range(h.ndim))
    pad = [(0, 0)] * h.ndim
    pad[axis] = (1, 1)
    hp = np.pad(h, pad)
    c = hp[idx(slice(1, 2 * n, 2))]
    left = hp[idx(slice(0, 2 * n - 1, 2))]
    right = hp[idx(slice(2, 2 * n + 1, 2))]
    return c + 0.5 * (left + right)


def interpolation_operator_3d(xH, Lc, L):
    """Trilinear prolongation: along x, then y, then z."""
    H = xH.reshape(Lc["nnz"], Lc["nny"], Lc["nnx"])
    h = _prolong_axis(_prolong_axis(_prolong_axis(H, 2), 1), 0)
    assert h.shape == (L["nnz"], L["nny"], L["nnx"])
    return h.ravel()


def restriction_operator_3d(xh, L, Lc):
    """R = P^T of the trilinear prolongation (unscaled, as the 2D R = I^T): along x, then y, then z."""
    h = xh.reshape(L["nnz"], L["nny"], L["nnx"])
    H = _restrict_axis(_restrict_axis(_restrict_axis(h, 2), 1), 0)
    assert H.shape == (Lc["nnz"], Lc["nny"], Lc["nnx"])
    return H.ravel()


def _restrict(x, L, Lc, periodic):
    if L.get("ndim") == 3:
        return restriction_operator_3d(x, L, Lc)
    return restriction_operator(x, L["nnx"], L["nny"], Lc["nnx"], Lc["nny"], periodic)


def _prolong(x, Lc, L, periodic):
    if L.get("ndim") == 3:
        return interpolation_operator_3d(x, Lc, L)
    return interpolation_operator(x, Lc["nnx"], Lc["nny"], L["nnx"], L["nny"], periodic)


def v_cycle(rhs, levels, tol, periodic, set_zeroP):
    """One V-cycle from x0 = 0.  multigrid.py:231-330.

    Down: pre-smooth = one Jacobi from zero (x = 2/3 r / L_diag), residual,
    restrict (pinned form zeroes gamma_p entries; Neumann form mean-projects
    r, x, r - Lx and R(.)).  Coarse: CG to ``tol`` (pinned) or one Jacobi
    from zero (Neumann).  Up: prolong, (Neumann: mean-project), add, one
    Jacobi against the stored, un-projected level residual."""
    N = len(levels)
    r_lv = [None] * N
    x_lv = [None] * N
    r_lv[0] = np.asarray(rhs, dtype=float)
    for n in range(N - 1):
        L, Lc = levels[n], levels[n + 1]
        if set_zeroP:
            g = L["gamma_p"]
            rh = r_lv[n].copy()
            rh[g] = 0
            xh = 2.0 / 3.0 * rh / L["L_diag"]
            xh[g] = 0
            Lx = emult.laplacian_element_mult(xh, L["Le"], L["nconn"], L["nn"])
            Lx[g] = 0
            rH = rh - Lx
            rH = _restrict(rH, L, Lc, periodic)
        else:
            rh = r_lv[n] - r_lv[n].mean()
            xh = 2.0 / 3.0 * rh / L["L_diag"]
            xh = xh - xh.mean()
            rH = rh - emult.laplacian_element_mult(xh, L["Le"], L["nconn"], L["nn"])
            rH = rH - rH.mean()
            rH = _restrict(rH, L, Lc, periodic)
            rH = rH - rH.mean()
        x_lv[n] = xh
        r_lv[n + 1] = rH
    C = levels[N - 1]
    if set_zeroP:
        g = C["gamma_p"]

        def op(x):
            y = emult.laplacian_element_mult(x, C["Le"], C["nconn"], C["nn"])
            y[g] = 0
            return y

        rN = r_lv[N - 1].copy()
        rN[g] = 0
        xN, _ = cg(op, rN, tol=tol, maxiter=1000)
        xN[g] = 0
    else:
        rN = r_lv[N - 1] - r_lv[N - 1].mean()
        xN = single_jacobi_iter(np.zeros(rN.shape[0]), rN, C["Le_off"], C["L_diag"], 0,
                                C["nconn"], C["nn"], False)
    x_lv[N - 1] = xN
    for n in range(N - 2, -1, -1):
        L, Lc = levels[n], levels[n + 1]
        xH = _prolong(x_lv[n + 1], Lc, L, periodic)
        if not set_zeroP:
            xH = xH - xH.mean()
        xh = x_lv[n] + xH
        x_lv[n] = single_jacobi_iter(xh, r_lv[n], L["Le_off"], L["L_diag"], L["gamma_p"],
                                     L["nconn"], L["nn"], set_zeroP)
    return x_lv[0]

