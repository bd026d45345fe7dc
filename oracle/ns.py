"""Navier-Stokes 4-step projection, host loop and passive scalar (oracle,
test-only).  Follows jax_fem_cfd/solvers/standard.py,
jax_fem_cfd/solvers/passive_transport.py and jax_fem_cfd/setup/simulation.py.
"""
from dataclasses import dataclass, field

import numpy as np

from . import elements as el
from . import emult
from . import mesh as ms
from . import multigrid as mg
from .solvers import bicgstab, cg


def solve_momentum(U_n, P_n, Me, Le, Ge, Ce_all_n, rho, mu, dt, nconn, nn, ndim, gamma_d, U_d,
                   F, tol, theta, gamma):
    """BiCGSTAB on (rho/dt) M_c + theta (rho C + mu L), Dirichlet rows -> identity.
    standard.py:14-43."""
    mask = None
    if gamma_d is not None:
        mask = np.zeros(ndim * nn, dtype=bool)
        mask[gamma_d] = True

    def lhs(U):
        res = ((rho / dt) * emult.consistent_mass_mult(U, Me, nconn, nn, ndim)
               + theta * emult.visc_conv_mult(U, Ce_all_n, Le, mu, rho, nconn, nn, ndim))
        if mask is not None:
            res = np.where(mask, U, res)
        return res

    rhs = ((rho / dt) * emult.consistent_mass_mult(U_n, Me, nconn, nn, ndim)
           - gamma * emult.gradient_element_mult(P_n, Ge, nconn, nn, ndim)
           - (1 - theta) * emult.visc_conv_mult(U_n, Ce_all_n, Le, mu, rho, nconn, nn, ndim) + F)
    if gamma_d is not None:
        rhs = rhs.copy()
        rhs[gamma_d] = U_d
        x0 = U_n.copy()
        x0[gamma_d] = U_d
    else:
        x0 = U_n
    return bicgstab(lhs, rhs, x0=x0, tol=tol)


def solve_pressure(U_star, P_n, Le, Ge, rho, dt, nconn, nn, ndim, gamma_p, F_p, set_zeroP, M_op,
                   tol, gamma):
    """CG on L (rows gamma_p zeroed when pinned), maxiter 1000.
    standard.py:85-125."""
    def lhs(P):
        res = emult.laplacian_element_mult(P, Le, nconn, nn)
        if set_zeroP:
            res[gamma_p] = 0
        return res

    rhs = (rho / dt) * (emult.divergence_element_mult(U_star, Ge, nconn, nn, ndim) - F_p)
    if gamma == 1:
        rhs = rhs + emult.laplacian_element_mult(P_n, Le, nconn, nn)
    if set_zeroP:
        rhs[gamma_p] = 0
        x0 = P_n.copy()
        x0[gamma_p] = 0
    else:
        rhs = rhs - rhs.mean()
        x0 = P_n
    P, it = cg(lhs, rhs, x0=x0, tol=tol, maxiter=1000, M=M_op)
    if set_zeroP:
        P = P.copy()
        P[gamma_p] = 0
    else:
        P = P - P.mean()
    return P, it


@dataclass
class Solver:
    """Constants assembled by setup_solver (standard.py:191-247)."""
    nn: int
    ndim: int
    nconn: np.ndarray
    M: np.ndarray
    Me: np.ndarray
    Le: np.ndarray
    Ge: np.ndarray
    wq: np.ndarray
    N: np.ndarray
    Nderiv: np.ndarray
    gamma_d: object
    gamma_p: object
    set_zeroP: bool
    tol: float
    periodic: bool
    precond: object
    M_op: object = None
    levels: list = field(default_factory=list)
    coords: np.ndarray = None
    bconn: object = None
    nvec: object = None
    L_diag: np.ndarray = None
    bshape: object = None


def setup_solver(coords, nconn, bconn, nvec, gamma_d, gamma_p, nn, mesh_func, num_elem, domain_size,
                 periodic, outlet_faces, solver_tol, set_zeroP, pressure_precond,
                 final_multigrid_mesh, partial_dirichlet=None, shape_func_ord=1, approx_keep=False):
    """standard.py:191-247: element constants from element 0 and the pressure
    preconditioner closure (jacobi | approx_inv | multigrid | None)."""
    ndim = len(num_elem)
    sf, bsf = el.shape_functions_for(ndim, shape_func_ord)
    wq, N, Nderiv = sf(coords[nconn[0]])
    M, Me = el.lumped_mass_matrix(nconn, wq, N, nn, ndim)
    Le = el.laplacian_element_calc(wq, Nderiv)
    Ge = el.gradient_element_calc(wq, N, Nderiv, ndim)
    s = Solver(nn=nn, ndim=ndim, nconn=nconn, M=M, Me=Me, Le=Le, Ge=Ge, wq=wq, N=N, Nderiv=Nderiv,
               gamma_d=gamma_d, gamma_p=gamma_p, set_zeroP=set_zeroP, tol=solver_tol,
               periodic=periodic, precond=pressure_precond, coords=coords, bconn=bconn, nvec=nvec,
               bshape=bsf)
    if pressure_precond == "jacobi":
        Ld = el.laplacian_diag(Le, nconn, nn)
        s.L_diag = Ld
        s.M_op = lambda r: r / Ld
    elif pressure_precond == "approx_inv":
        Ld = el.laplacian_diag(Le, nconn, nn)
        s.L_diag = Ld

        def approx_inv(r):
            # standard.py:228-234: the gamma_p rows of the two inner products are set to 0.
            # The passive solver's own closure (passive_transport.py:154-162, approx_keep) sets
            # them to D1x and D1L instead.
            D1x = r / Ld
            D1L = emult.laplacian_element_mult(D1x, Le, nconn, nn) / Ld
            D1L[gamma_p] = D1x[gamma_p] if approx_keep else 0
            L2 = emult.laplacian_element_mult(D1L, Le, nconn, nn)
            L2[gamma_p] = D1L[gamma_p] if approx_keep else 0
            return 3 * D1x - 3 * D1L + L2 / Ld
        s.M_op = approx_inv
    elif pressure_precond == "multigrid":
        levels = mg.multigrid_setup(num_elem, final_multigrid_mesh, domain_size, outlet_faces,
                                    mesh_func, sf, partial_dirichlet, shape_func_ord)
        s.levels = levels
        s.M_op = lambda r: mg.v_cycle(r, levels, solver_tol, periodic, set_zeroP)
    return s


def four_step_timestep(s, U_n, P_n, dt, rho, mu, U_d, F_p, F, theta=0.5):
    """Choi et al. 4-step projection.  standard.py:148-165."""
    nn, ndim, nconn = s.nn, s.ndim, s.nconn
    Ce = el.conservative_convec_elem_calc(nconn, U_n, s.wq, s.N, s.Nderiv, nn, ndim, 0.5)
    U_hat, U_it = solve_momentum(U_n, P_n, s.Me, s.Le, s.Ge, Ce, rho, mu, dt, nconn, nn, ndim,
                                 s.gamma_d, U_d, F, s.tol, theta, 1)
    U_star = U_hat + (dt / rho) * emult.gradient_element_mult(P_n, s.Ge, nconn, nn, ndim) / s.M
    P, P_it = solve_pressure(U_star, P_n, s.Le, s.Ge, rho, dt, nconn, nn, ndim, s.gamma_p, F_p,
                             s.set_zeroP, s.M_op, s.tol, 0)
    U = U_star - (dt / rho) * emult.gradient_element_mult(P, s.Ge, nconn, nn, ndim) / s.M
    if s.gamma_d is not None:
        U = U.copy()
        U[s.gamma_d] = U_d
    return U, P, U_it, P_it


def compute(s, U_d, S_func, t_n, t):
    """F_p (boundary integral of q u.n over Dirichlet edges) and
    F = M_c (S(t_n) + S(t)) / 2.  standard.py:167-188."""
    F, F_p = 0, 0
    if s.gamma_d is not None and s.ndim == 2:
        F_p = el.u_dot_n_boundary_integral_1d(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords,
                                              s.bshape or el.lin2_shape_functions)
    elif s.gamma_d is not None and s.nconn.shape[1] == 27:
        F_p = el.u_dot_n_boundary_integral_quad9_faces(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords)
    elif s.gamma_d is not None:
        F_p = el.u_dot_n_boundary_integral_2d(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords)
    if S_func is not None:
        S = (S_func(t_n) + S_func(t)) / 2
        F = emult.consistent_mass_mult(S, s.Me, s.nconn, s.nn, s.ndim)
    return F_p, F


def num_steps_for(t_final, dt):
    """simulation.py:258-260."""
    n = int(t_final / dt)
    if n * dt < t_final:
        n += 1
    return n


def timestep_loop(s, t_final, dt, save_interval, rho, mu, U_d, U0, P0, S_func=None):
    """Host loop: per step compute -> timestep; record iteration counts and
    solutions on save steps.  simulation.py:229-360 (printing and HDF5
    streaming left out)."""
    U, P = U0, P0
    U_it_l, P_it_l, steps, U_list = [], [], [], [np.array(U)]
    n = num_steps_for(t_final, dt)
    for step in range(1, n + 1):
        t_n, t = (step - 1) * dt, step * dt
        F_p, F = compute(s, U_d, S_func, t_n, t)
        U, P, ui, pi = four_step_timestep(s, U, P, dt, rho, mu, U_d, F_p, F)
        if step % save_interval == 0 or step == n:
            U_list.append(np.array(U))
            U_it_l.append(ui)
            P_it_l.append(pi)
            steps.append(step)
    return U, P, np.array(U_it_l), np.array(P_it_l), np.array(steps), np.stack(U_list)


# --------------------------------------------------------------------------------------------
# Reference configurations used by the tests and the CPU baseline
# --------------------------------------------------------------------------------------------

def build(num_elem, domain_size, periodic, dirichlet_faces, outlet_faces, solver_tol, set_zeroP,
          pressure_precond, final_multigrid_mesh, partial_dirichlet=None, shape_func_ord=1, passive=False):
    """simulation.setup for the 'standard' solver (simulation.py:185-200); passive = True builds
    the pressure preconditioner of the passive-transport solver (passive_transport.py:133-172)."""
    coords, nconn, surf, nn, h, bconn, nvec, gamma_d, U_d, gamma_p, mesh_func = ms.simple_mesh(
        num_elem, domain_size, periodic, dirichlet_faces, outlet_faces, partial_dirichlet, shape_func_ord)
    s = setup_solver(coords, nconn, bconn, nvec, gamma_d, gamma_p, nn, mesh_func, num_elem,
                     domain_size, periodic, outlet_faces, solver_tol, set_zeroP, pressure_precond,
                     final_multigrid_mesh, partial_dirichlet, shape_func_ord, approx_keep=passive)
    return s, surf, h, U_d


def taylor_green(coords, t, rho, mu):
    """Analytic Taylor-Green vortex.  examples/taylor_green.py:6-14."""
    x, y = coords[:, 0], coords[:, 1]
    u = np.sin(x) * np.cos(y) * np.exp((-2 * mu / rho) * t)
    v = -np.cos(x) * np.sin(y) * np.exp((-2 * mu / rho) * t)
    p = (rho / 4) * (np.cos(2 * x) + np.cos(2 * y)) * np.exp((-4 * mu / rho) * t)
    return u, v, p


def cavity_lid(s, surf, U_d, u_top=1.0):
    """Lid u = u_top on the top face without its corners.  cavity.py:115-117."""
    return ms.set_face_velocity(U_d, surf[2][1:-1], s.gamma_d, u_top, s.nn, 1)


def channel_inflow(s, surf, U_d, ly, lz, u_max=1.0):
    """3D channel: parabolic inflow u = 16 u_max (y/ly)(1-y/ly)(z/lz)(1-z/lz) on the left
    face (surfnodes[3]); the profile vanishes on the face edges, matching the no-slip walls."""
    nodes = surf[3]
    y, z = s.coords[nodes, 1] / ly, s.coords[nodes, 2] / lz
    return ms.set_face_velocity(U_d, nodes, s.gamma_d, 16 * u_max * y * (1 - y) * z * (1 - z), s.nn, 1)


# --------------------------------------------------------------------------------------------
# Passive scalar transport (passive_transport.py)
# --------------------------------------------------------------------------------------------

def update_phi(phi_n, nconn, Me, Ce_all, Ce_all_n, Le, D, dt, nn, gamma_d_phi, phi_d, tol):
    """Crank-Nicolson ADE by BiCGSTAB.  passive_transport.py:39-61."""
    mask = None
    if gamma_d_phi is not None:
        mask = np.zeros(nn, dtype=bool)
        mask[gamma_d_phi] = True

    def lhs(phi):
        res = ((1 / dt) * emult.consistent_mass_mult(phi, Me, nconn, nn, 1)
               + 0.5 * emult.scalar_visc_conv_mult(phi, Ce_all, Le, D, nconn, nn))
        if mask is not None:
            res = np.where(mask, phi, res)
        return res

    rhs = ((1 / dt) * emult.consistent_mass_mult(phi_n, Me, nconn, nn, 1)
           - 0.5 * emult.scalar_visc_conv_mult(phi_n, Ce_all_n, Le, D, nconn, nn))
    if gamma_d_phi is not None:
        rhs = rhs.copy()
        rhs[gamma_d_phi] = phi_d
        x0 = phi_n.copy()
        x0[gamma_d_phi] = phi_d
    else:
        x0 = phi_n
    return bicgstab(lhs, rhs, x0=x0, tol=tol)


def update_phi_SUPG(phi_n, nconn, Mtau_all_n, Ctau_all_n, Me, Ce_all, Ce_all_n, Le, D, dt, nn,
                    gamma_d_phi, phi_d, tol):
    """SUPG-stabilised CN ADE by BiCGSTAB.  passive_transport.py:14-37."""
    mask = None
    if gamma_d_phi is not None:
        mask = np.zeros(nn, dtype=bool)
        mask[gamma_d_phi] = True

    def lhs(phi):
        res = ((1 / dt) * emult.scalar_consistent_mass_SUPG_mult(phi, Mtau_all_n, Me, nconn, nn)
               + 0.5 * emult.scalar_stabilized_visc_conv_mult(phi, Ce_all, Ctau_all_n, Le, D,
                                                              nconn, nn))
        if mask is not None:
            res = np.where(mask, phi, res)
        return res

    rhs = ((1 / dt) * emult.scalar_consistent_mass_SUPG_mult(phi_n, Mtau_all_n, Me, nconn, nn)
           - 0.5 * emult.scalar_stabilized_visc_conv_mult(phi_n, Ce_all_n, Ctau_all_n, Le, D,
                                                          nconn, nn))
    if gamma_d_phi is not None:
        rhs = rhs.copy()
        rhs[gamma_d_phi] = phi_d
        x0 = phi_n.copy()
        x0[gamma_d_phi] = phi_d
    else:
        x0 = phi_n
    return bicgstab(lhs, rhs, x0=x0, tol=tol)


def passive_timestep(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi, supg):
    """NS step, then Ce(U_n, beta=1) and Ce(U, beta=1) and the scalar update.
    passive_transport.py:64-97."""
    U, P, U_it, P_it = four_step_timestep(s, U_n, P_n, dt, rho, mu, U_d, F_p, F)
    phi, phi_it = scalar_step(s, U_n, U, phi_n, dt, D, phi_d, gamma_d_phi, supg)
    return U, P, phi, U_it, P_it, phi_it


def scalar_step(s, U_n, U, phi_n, dt, D, phi_d, gamma_d_phi, supg):
    """The scalar half of the passive step for given old / new velocities
    (passive_transport.py:80-95): convection matrices with beta = 1, SUPG tau and
    Ct from U_n (tau's |Ce| from the new-velocity Ce), then the CN solve."""
    nconn, nn, ndim = s.nconn, s.nn, s.ndim
    Ce_n = el.conservative_convec_elem_calc(nconn, U_n, s.wq, s.N, s.Nderiv, nn, ndim, 1)
    Ce = el.conservative_convec_elem_calc(nconn, U, s.wq, s.N, s.Nderiv, nn, ndim, 1)
    if supg:
        Ct_n = el.Ctau_tilde_elem_calc(nconn, U_n, s.wq, s.N, s.Nderiv, nn, ndim)
        tau_n = el.calculate_tau(nconn, U_n, Ce, Ct_n, s.N, nn, ndim, D, dt, r=2)
        Ctau_n = tau_n[:, None, None] * Ct_n
        Mtau_n = tau_n[:, None, None] * Ce_n.transpose(0, 2, 1)
        phi, phi_it = update_phi_SUPG(phi_n, nconn, Mtau_n, Ctau_n, s.Me, Ce, Ce_n, s.Le, D, dt,
                                      nn, gamma_d_phi, phi_d, s.tol)
    else:
        phi, phi_it = update_phi(phi_n, nconn, s.Me, Ce, Ce_n, s.Le, D, dt, nn, gamma_d_phi,
                                 phi_d, s.tol)
    return phi, phi_it


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))

