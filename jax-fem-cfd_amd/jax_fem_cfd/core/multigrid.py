"""Geometric multigrid preconditioner on the GPU.

``multigrid_setup`` returns the reference's tuple
``(lvl_nums, Le_lvls, nconn_lvls, L_diag_lvls, gamma_p_lvls)``
(jax_fem_cfd/core/multigrid.py:9-88) and ``v_cycle`` runs one V-cycle from
zero in the HIP kernels of libfcfd (:231-330).  The hierarchy is rebuilt on the
device once per distinct configuration and cached.
"""
import numpy as np

from .. import _native as nat
from .._context import Context, _h_from_shape


def multigrid_setup(num_elem, num_elemf, domain_size, outlet_faces, mesh_func, shapefunc, shape_func_ord,
                    partial_dirichlet):
    """multigrid.py:9-88: per level mesh, Le, L_diag, Le_off and gamma_p; level
    count from the x ratio only."""
    if shape_func_ord != 1:
        raise NotImplementedError("multigrid is built for Q1 elements")
    nex, ney = num_elem
    N = int(np.log2(nex // num_elemf[0])) + 1
    lvl_nums = np.zeros((N + 1, 5), dtype=np.int64)
    Le_lvls = np.zeros((N, 2, 4, 4))
    nconn_lvls, L_diag_lvls, gamma_p_lvls = [], [], []
    from ..core.element_calc import laplacian_element_calc
    for i in range(N):
        node_coords, nconn, surfnodes, bconns, _ = mesh_func(nex, ney, *domain_size)
        wq, _, Nderiv = shapefunc(node_coords[nconn[0]])
        nn = node_coords.shape[0]
        nnx, nny = nex + 1, ney + 1
        outlet = []
        for f in outlet_faces or []:
            outlet.append(surfnodes[f])
        if partial_dirichlet:
            for f, (s_frac, e_frac, rem) in partial_dirichlet.items():
                nfe = bconns[f].shape[0]
                s_idx, e_idx = round(s_frac * nfe), round(e_frac * nfe)
                if rem == "outlet":
                    outlet.append(surfnodes[f][:s_idx])
                    outlet.append(surfnodes[f][e_idx + 1:])
        gamma_p = np.unique(np.concatenate(outlet)).astype(np.int64) if outlet else np.array([0])
        Le = laplacian_element_calc(wq, Nderiv)
        # L_diag on the host: set-up constant, scatter of diag(Le) (element_calc.py:99-109)
        L_diag = np.bincount(nconn.ravel(), np.tile(np.diag(Le), nconn.shape[0]), minlength=nn)
        lvl_nums[i + 1] = [lvl_nums[i, 0] + nn, nnx, nny, lvl_nums[i, 3] + nex * ney,
                           lvl_nums[i, 4] + gamma_p.shape[0]]
        Le_lvls[i, 0] = Le
        Le_lvls[i, 1] = Le - np.diag(np.diag(Le))
        nconn_lvls.append(nconn)
        L_diag_lvls.append(L_diag)
        gamma_p_lvls.append(gamma_p)
        nex, ney = nex // 2, ney // 2
    lvl_nums = tuple(tuple(int(x) for x in row) for row in lvl_nums)
    return lvl_nums, Le_lvls, np.concatenate(nconn_lvls), np.concatenate(L_diag_lvls), np.concatenate(gamma_p_lvls)


_mg_cache = {}


def _mg_context(lvl_nums, nconn_lvls, gamma_p_lvls, tol, periodic, set_zeroP, hx, hy, precision="fp64"):
    N = len(lvl_nums) - 1
    nnx, nny = lvl_nums[1][1], lvl_nums[1][2]
    nex, ney = nnx - 1, nny - 1
    gp = [np.asarray(gamma_p_lvls[lvl_nums[i][4]:lvl_nums[i + 1][4]]) for i in range(N)]
    key = (lvl_nums, float(tol), bool(periodic), bool(set_zeroP), round(hx, 15), round(hy, 15), precision,
           tuple(g.tobytes() for g in gp))
    ctx = _mg_cache.get(key)
    if ctx is None:
        ctx = Context(nex, ney, hx * nex, hy * ney, periodic, tol=tol, set_zeroP=set_zeroP, precond="multigrid",
                      mg_levels=N, precision=precision, gamma_p=gp[0], mg_gamma_p=gp)
        _mg_cache[key] = ctx
    return ctx


def v_cycle(rhs, lvl_nums, Le_lvls, nconn_lvls, L_diag_lvls, gamma_p_lvls, tol, periodic, set_zeroP):
    """multigrid.py:231-330: one V-cycle from x0 = 0 on the GPU."""
    Le = nat.to_numpy(Le_lvls)[0, 0]
    # element size from Le-independent data: the fine level spans lvl_nums[1] nodes of a
    # uniform grid; h only enters the (h-independent in 2D) Laplacian, so unit cells are used
    ctx = _mg_context(lvl_nums, nconn_lvls, nat.to_numpy(gamma_p_lvls), tol, periodic, set_zeroP, 1.0, 1.0)
    if not np.allclose(ctx.element_constant("Le"), Le, rtol=0, atol=1e-14):
        raise ValueError("Le_lvls does not match the Q1 Laplacian of a uniform square grid")
    return ctx.vcycle(rhs)


_transfer_cache = {}


def _transfer_context(nnh_x, nnh_y, periodic):
    """A two-level context whose level 0 is the fine node grid of the reference's transfer call
    (nnh_x x nnh_y nodes; periodic grids count the padded wrap column and row, as the reference's
    arguments do) and whose level 1 is its coarse grid.  Unit cells: the transfers do not depend
    on the element size."""
    nex, ney = int(nnh_x) - 1, int(nnh_y) - 1
    if nex < 2 or ney < 2 or nex % 2 or ney % 2:
        raise ValueError("the fine grid must have an even number of elements per direction")
    key = (nex, ney, bool(periodic))
    ctx = _transfer_cache.get(key)
    if ctx is None:
        ctx = Context(nex, ney, float(nex), float(ney), bool(periodic), precond="multigrid", mg_levels=2)
        _transfer_cache[key] = ctx
    return ctx


def interpolation_operator(xH, nnH_x, nnH_y, nnh_x, nnh_y, periodic):
    """multigrid.py:90-140: bilinear interpolation of xH (coarse node grid nnH_x x nnH_y) to the
    fine grid nnh_x x nnh_y with half the element size, on the GPU (fcfd_interpolate)."""
    if (int(nnh_x) - 1) != 2 * (int(nnH_x) - 1) or (int(nnh_y) - 1) != 2 * (int(nnH_y) - 1):
        raise ValueError("the fine grid must have twice the coarse grid's elements per direction")
    return _transfer_context(nnh_x, nnh_y, periodic).interpolate(xH, 0)


def restriction_operator(xh, nnh_x, nnh_y, nnH_x, nnH_y, periodic):
    """multigrid.py:142-212: R = I^T (no 1/4 full-weighting factor) from the fine grid
    nnh_x x nnh_y to the coarse grid nnH_x x nnH_y, on the GPU (fcfd_restrict)."""
    if (int(nnh_x) - 1) != 2 * (int(nnH_x) - 1) or (int(nnh_y) - 1) != 2 * (int(nnH_y) - 1):
        raise ValueError("the fine grid must have twice the coarse grid's elements per direction")
    return _transfer_context(nnh_x, nnh_y, periodic).restrict(xh, 0)


def single_jacobi_iter(x, rhs, Le_off, L_diag, gamma_p, nconn, nn, set_zeroP):
    """multigrid.py:214-228: one damped Jacobi sweep (omega = 2/3) with the given Le_off and
    L_diag on the grid behind nconn; gamma_p entries zeroed before and after when set_zeroP,
    the mean subtracted otherwise (fcfd_jacobi_sweep)."""
    from .._context import grid_context
    ctx = grid_context(nconn, int(nn))
    return ctx.jacobi_sweep(x, rhs, Le_off, L_diag, gamma_p if set_zeroP else None, set_zeroP)
