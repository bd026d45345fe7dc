"""Incompressible Newtonian Navier-Stokes, 4-step pressure projection.

``setup_solver`` has the signature and return value of
jax_fem_cfd/solvers/standard.py:191-267: it builds one native context (element
constants, lumped mass, Dirichlet masks, multigrid hierarchy, workspaces) and
returns the closures ``compute(U_d, S_func, t_n, t) -> (F_p, F)`` and
``timestep(U_n, P_n, dt, rho, mu, U_d, F_p, F) -> (U, P, U_iters, P_iters)``.
One timestep = one call of fcfd_four_step_timestep: Ce build, BiCGSTAB
momentum solve, U*, MG-PCG pressure solve and projection, all in HIP kernels
(standard.py:148-165).
"""
import numpy as np

from .. import _native as nat
from .._context import Context
from ..setup.simulation import Config


def _mg_levels(num_elem, final):
    return int(np.log2(num_elem[0] // final[0])) + 1


def _mg_pins(num_elem, domain_size, outlet_faces, partial_dirichlet, mesh_func, nlev, shape_func_ord=1):
    """per-level gamma_p as multigrid_setup builds them (multigrid.py:35-62)."""
    pins = []
    if len(num_elem) == 3:
        # box-grid hierarchy: the outlet faces of each level's hex8 mesh
        nex, ney, nez = num_elem
        for _ in range(nlev):
            _, _, surf, _, _ = mesh_func(nex, ney, nez, *domain_size)
            outlet = [surf[f] for f in outlet_faces or []]
            pins.append(np.unique(np.concatenate(outlet)).astype(np.int64) if outlet
                        else np.array([0], dtype=np.int64))
            nex, ney, nez = nex // 2, ney // 2, nez // 2
        return pins
    nex, ney = num_elem
    for _ in range(nlev):
        _, _, surf, bconns, _ = mesh_func(nex, ney, *domain_size)
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
        pins.append(np.unique(np.concatenate(outlet)).astype(np.int64) if outlet else np.array([0], dtype=np.int64))
        nex, ney = nex // 2, ney // 2
    return pins


def build_context(nconn, bconn, nvec, gamma_d, gamma_p, nn, mesh_func, config: Config, ns_config=None,
                  gamma_d_phi=None, passive=False):
    """One native context for the configuration.  passive: the pressure preconditioner of the
    passive-transport solver, whose approx_inv closure keeps D1x and D1L at gamma_p
    (passive_transport.py:154-162) where standard.py:225-234 zeroes them."""
    mc = config.mesh_config
    sc = ns_config if ns_config is not None else config.solver_config
    ce_format = sc.ce_format or ("f64" if sc.precision == "fp64" else "f32")
    if mc.shape_func_ord == 2 and len(mc.num_elem) == 3:
        if gamma_d_phi is not None:
            raise NotImplementedError("the 3D GPU path has no passive scalar")
        nex, ney, nez = mc.num_elem
        nlev = _mg_levels(mc.num_elem, sc.final_multigrid_mesh) if sc.pressure_precond == "multigrid" else 1
        mg_pins = None
        if sc.pressure_precond == "multigrid" and sc.set_zeroP:
            mg_pins = _mg_pins(mc.num_elem, mc.domain_size, sc.outlet_faces, None, mesh_func, nlev)
        return Context(nex, ney, mc.domain_size[0], mc.domain_size[1], False, tol=sc.solver_tol,
                       set_zeroP=sc.set_zeroP, precond=sc.pressure_precond, mg_levels=nlev, precision=sc.precision,
                       ce_format=ce_format, gamma_d=gamma_d,
                       gamma_p=np.atleast_1d(gamma_p) if (sc.set_zeroP or sc.pressure_precond == "approx_inv")
                       else None, mg_gamma_p=mg_pins, bconn=bconn, nvec=nvec, nez=nez, lz=mc.domain_size[2],
                       order=2)
    if mc.shape_func_ord == 2:
        nex, ney = mc.num_elem
        nlev = _mg_levels(mc.num_elem, sc.final_multigrid_mesh) if sc.pressure_precond == "multigrid" else 1
        mg_pins = None
        if sc.pressure_precond == "multigrid" and sc.set_zeroP:
            mg_pins = _mg_pins(mc.num_elem, mc.domain_size, sc.outlet_faces, mc.partial_dirichlet, mesh_func, nlev, 2)
        return Context(nex, ney, mc.domain_size[0], mc.domain_size[1], mc.periodic, tol=sc.solver_tol,
                       set_zeroP=sc.set_zeroP, precond=sc.pressure_precond, mg_levels=nlev, precision=sc.precision,
                       ce_format=ce_format, gamma_d=gamma_d,
                       gamma_p=np.atleast_1d(gamma_p) if (sc.set_zeroP or sc.pressure_precond == "approx_inv")
                       else None, mg_gamma_p=mg_pins, bconn=bconn, nvec=nvec, order=2, gamma_d_phi=gamma_d_phi,
                       approx_inv_variant=1 if passive else 0)
    if mc.shape_func_ord != 1:
        raise NotImplementedError("the GPU timestep is built for quad4, quad9, hex8 and hex27 meshes")
    if len(mc.num_elem) == 3:
        if gamma_d_phi is not None:
            raise NotImplementedError("the 3D GPU path has no passive scalar")
        nex, ney, nez = mc.num_elem
        nlev = _mg_levels(mc.num_elem, sc.final_multigrid_mesh) if sc.pressure_precond == "multigrid" else 1
        mg_pins = None
        if sc.pressure_precond == "multigrid" and sc.set_zeroP:
            mg_pins = _mg_pins(mc.num_elem, mc.domain_size, sc.outlet_faces, None, mesh_func, nlev)
        return Context(nex, ney, mc.domain_size[0], mc.domain_size[1], False, tol=sc.solver_tol,
                       set_zeroP=sc.set_zeroP, precond=sc.pressure_precond, mg_levels=nlev, precision=sc.precision,
                       ce_format=ce_format, gamma_d=gamma_d,
                       gamma_p=np.atleast_1d(gamma_p) if (sc.set_zeroP or sc.pressure_precond == "approx_inv")
                       else None, mg_gamma_p=mg_pins, bconn=bconn, nvec=nvec, nez=nez, lz=mc.domain_size[2])
    nex, ney = mc.num_elem
    nlev = _mg_levels(mc.num_elem, sc.final_multigrid_mesh) if sc.pressure_precond == "multigrid" else 1
    mg_pins = None
    if sc.pressure_precond == "multigrid" and sc.set_zeroP:
        mg_pins = _mg_pins(mc.num_elem, mc.domain_size, sc.outlet_faces, mc.partial_dirichlet, mesh_func, nlev)
    return Context(nex, ney, mc.domain_size[0], mc.domain_size[1], mc.periodic, tol=sc.solver_tol,
                   set_zeroP=sc.set_zeroP, precond=sc.pressure_precond, mg_levels=nlev, precision=sc.precision,
                   ce_format=ce_format, gamma_d=gamma_d,
                   gamma_p=np.atleast_1d(gamma_p) if (sc.set_zeroP or sc.pressure_precond == "approx_inv") else None,
                   mg_gamma_p=mg_pins, bconn=bconn, nvec=nvec, gamma_d_phi=gamma_d_phi,
                   approx_inv_variant=1 if passive else 0)


def setup_solver(node_coords, nconn, bconn, nvec, gamma_d, gamma_p, nn, mesh_func, config: Config):
    """standard.py:191-267."""
    ctx = build_context(nconn, bconn, nvec, gamma_d, gamma_p, nn, mesh_func, config)
    has_dirichlet = gamma_d is not None
    Me = ctx.element_constant("Me")

    def compute_loaded(U_d, S_func, t_n, t):
        """standard.py:167-188: F_p = int q (U_d . n) (HIP kernel), F = M_c (S(t_n) + S(t)) / 2."""
        F_p, F = 0, 0
        if has_dirichlet:
            F_p = ctx.boundary_integral(U_d)
        if S_func is not None:
            S = (nat.as_dev(S_func(t_n)) + nat.as_dev(S_func(t))) / 2
            if ctx.ndim == 3 or ctx.order == 2:
                F = ctx.mass_matvec(S)
            else:
                F = ctx._empty(2 * nn)
                for d in range(2):
                    ctx.elem_mult_const(Me, S[d * nn:(d + 1) * nn], out=F[d * nn:(d + 1) * nn])
        return F_p, F

    def timestep_loaded(U_n, P_n, dt, rho, mu, U_d, F_p, F):
        """standard.py:256-265 -> fcfd_four_step_timestep."""
        return ctx.four_step_timestep(U_n, P_n, dt, rho, mu, U_d if has_dirichlet else None, F_p, F)

    timestep_loaded.context = ctx
    compute_loaded.context = ctx
    return compute_loaded, timestep_loaded
