"""Navier-Stokes step plus passive scalar transport.

``setup_solver`` has the signature and return value of
jax_fem_cfd/solvers/passive_transport.py:100-180 and returns
``compute(U_d, S_func, t_n, t) -> (F_p, F)`` and
``timestep(U_n, P_n, phi_n, dt, rho, mu, D, U_d, phi_d, F_p, F, phi_S_func)
-> (U, P, phi, U_iters, P_iters, phi_iters)``.  One call = one
fcfd_passive_timestep: the 4-step projection, then Ce(U_n), Ce(U) (beta = 1),
optional SUPG tau / Ct and the Crank-Nicolson BiCGSTAB of the scalar
(passive_transport.py:64-97), all in HIP kernels.  As in the reference the
multigrid hierarchy uses the non-periodic mesh (:107-122) and phi_S_func is
accepted but not used by the scalar update.
"""
import numpy as np

from .. import _native as nat
from ..setup import regular_mesh as rm
from .standard import build_context


def setup_solver(node_coords, nconn, bconn, nvec, gamma_d, gamma_p, nn, gamma_d_phi, mesh_func, config):
    """passive_transport.py:100-180."""
    ns = config.solver_config.NS_config
    ade = config.solver_config.ADE_config
    if config.mesh_config.periodic:
        raise NotImplementedError("the passive-scalar solver uses non-periodic meshes (as the reference)")
    # the mesh function follows the element order (passive_transport.py:107-116)
    mesh_func = rm.create_mesh_2d_quad9 if config.mesh_config.shape_func_ord == 2 else rm.create_mesh_2d
    ctx = build_context(nconn, bconn, nvec, gamma_d, gamma_p, nn, mesh_func, config, ns_config=ns,
                        gamma_d_phi=gamma_d_phi, passive=True)
    has_dirichlet = gamma_d is not None
    Me = ctx.element_constant("Me")
    supg = bool(ade.phi_SUPG)

    def compute_loaded(U_d, S_func, t_n, t):
        F_p, F = 0, 0
        if has_dirichlet:
            F_p = ctx.boundary_integral(U_d)
        if S_func is not None:
            S = (nat.as_dev(S_func(t_n)) + nat.as_dev(S_func(t))) / 2
            if ctx.order == 2:
                F = ctx.mass_matvec(S)
            else:
                F = ctx._empty(2 * nn)
                for d in range(2):
                    ctx.elem_mult_const(Me, S[d * nn:(d + 1) * nn], out=F[d * nn:(d + 1) * nn])
        return F_p, F

    def timestep_loaded(U_n, P_n, phi_n, dt, rho, mu, D, U_d, phi_d, F_p, F, phi_S_func):
        return ctx.passive_timestep(U_n, P_n, phi_n, dt, rho, mu, D, U_d if has_dirichlet else None,
                                    None if gamma_d_phi is None else phi_d, F_p, F, supg)

    timestep_loaded.context = ctx
    timestep_loaded.rollout_info = dict(supg=supg, has_dirichlet=has_dirichlet,
                                        has_phi_dirichlet=gamma_d_phi is not None)
    compute_loaded.context = ctx
    return compute_loaded, timestep_loaded


def _ade_solve(lhs_K0, lhs_mats, rhs_K0, rhs_mats, ctx, phi_n, gamma_d_phi, phi_d, tol):
    """BiCGSTAB (the reference's, on the device vectors) on a scalar operator given as a shared
    element matrix K0 plus weighted per-element matrices, Dirichlet rows replaced by the identity
    (passive_transport.py:17-37, :41-61)."""
    from ..utils.jax_iterative_solvers import bicgstab
    t = nat.torch()
    mask = None
    if gamma_d_phi is not None:
        mask = t.zeros(ctx.nn, dtype=t.bool, device=nat.device())
        mask[nat.as_index(gamma_d_phi)] = True

    def apply(K0, mats, x):
        (A1, c1), (A2, c2) = (mats + [(None, 0.0), (None, 0.0)])[:2]
        return ctx.elem_mult_stored(x, K0=K0, A1=A1, c1=c1, A2=A2, c2=c2)

    def lhs(x):
        y = apply(lhs_K0, lhs_mats, x)
        return y if mask is None else t.where(mask, nat.as_dev(x), y)

    phi_n = nat.as_dev(phi_n)
    rhs = apply(rhs_K0, rhs_mats, phi_n)
    x0 = phi_n.clone()
    if mask is not None:
        gd = nat.as_index(gamma_d_phi)
        rhs[gd] = nat.as_dev(phi_d) if not np.isscalar(phi_d) else float(phi_d)
        x0[gd] = rhs[gd]
    return bicgstab(lhs, rhs, x0=x0, tol=tol)


def _grid_ctx(nconn, nn):
    """The grid behind nconn (the element sizes do not enter: every operator is passed in)."""
    from .._context import grid_context
    return grid_context(nconn, int(nn))


def update_phi(phi_n, nconn, Me, Ce_all, Ce_all_n, Le, D, dt, nn, gamma_d_phi, phi_d, tol):
    """passive_transport.py:39-61 with its argument list: Crank-Nicolson advection-diffusion,
    [(1/dt) M_c + (C + D L) / 2] phi = (1/dt) M_c phi_n - (C_n + D L) phi_n / 2 with the given
    per-element convection matrices (ne, 4, 4); returns (phi, iters)."""
    ctx = _grid_ctx(nconn, nn)
    Me, Le = np.asarray(nat.to_numpy(Me)), np.asarray(nat.to_numpy(Le))
    return _ade_solve(Me / dt + 0.5 * D * Le, [(Ce_all, 0.5)], Me / dt - 0.5 * D * Le, [(Ce_all_n, -0.5)], ctx,
                      phi_n, gamma_d_phi, phi_d, tol)


def update_phi_SUPG(phi_n, nconn, Mtau_all_n, Ctau_all_n, Me, Ce_all, Ce_all_n, Le, D, dt, nn, gamma_d_phi, phi_d,
                    tol):
    """passive_transport.py:14-37 with its argument list: the same step with the SUPG mass
    (Me + Mtau_e) and streamline term Ctau_e; returns (phi, iters)."""
    ctx = _grid_ctx(nconn, nn)
    Me, Le = np.asarray(nat.to_numpy(Me)), np.asarray(nat.to_numpy(Le))
    Mt, Ct = nat.as_dev(Mtau_all_n), nat.as_dev(Ctau_all_n)
    CS, CSn = nat.as_dev(Ce_all) + Ct, nat.as_dev(Ce_all_n) + Ct
    return _ade_solve(Me / dt + 0.5 * D * Le, [(Mt, 1.0 / dt), (CS, 0.5)], Me / dt - 0.5 * D * Le,
                      [(Mt, 1.0 / dt), (CSn, -0.5)], ctx, phi_n, gamma_d_phi, phi_d, tol)


def timestep(U_n, P_n, M, Le, Ge, rho, mu, dt, nconn, wq, N, Nderiv, gamma_d, U_d, gamma_p, F_p, nn, ndim, stab,
             p_precond, p_M_op, tol, set_zeroP, F, phi_n, gamma_d_phi, phi_d, D, Me, phi_S_func, phi_SUPG):
    """passive_transport.py:64-97 with its argument list: the four-step NS step, then the scalar
    step (plain or SUPG) on Ce(U_n), Ce(U) with beta = 1, in one native call
    (fcfd_passive_timestep); p_M_op as in standard.solve_pressure.  Returns
    (U, P, phi, U_iters, P_iters, phi_iters)."""
    from .standard import _op_context, _sizes_from_Ge, _check_const
    if ndim != 2:
        raise NotImplementedError("the passive-scalar step is built for 2D grids")
    if stab:
        raise NotImplementedError("the stabilised momentum variant (solve_momentum_stab) is not on this path")
    hx, hy = _sizes_from_Ge(Ge)
    ctx = _op_context(nconn, nn, hx, hy, gamma_d=gamma_d, gamma_p=gamma_p, set_zeroP=set_zeroP, precond=p_precond,
                      M_op=p_M_op, tol=tol, gamma_d_phi=gamma_d_phi, passive=True)
    _check_const(ctx, "Me", Me)
    _check_const(ctx, "Le", Le)
    Ud = None if gamma_d is None else nat.as_dev(np.asarray(nat.to_numpy(U_d), dtype=np.float64))
    return ctx.passive_timestep(U_n, P_n, phi_n, dt, rho, mu, D, Ud, None if gamma_d_phi is None else phi_d, F_p, F,
                                bool(phi_SUPG))
