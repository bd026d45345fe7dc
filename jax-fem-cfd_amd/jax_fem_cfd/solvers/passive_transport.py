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


def update_phi(phi_n, U_n, U, dt, D, ctx, phi_d=None):
    """update_phi (passive_transport.py:39-61) on a context's grid: Crank-Nicolson
    advection-diffusion with Ce(U_n), Ce(U) (beta = 1); returns (phi, iters)."""
    return ctx.solve_scalar(U_n, U, phi_n, dt, D, phi_d, supg=False)


def update_phi_SUPG(phi_n, U_n, U, dt, D, ctx, phi_d=None):
    """update_phi_SUPG (passive_transport.py:14-37) on a context's grid."""
    return ctx.solve_scalar(U_n, U, phi_n, dt, D, phi_d, supg=True)
