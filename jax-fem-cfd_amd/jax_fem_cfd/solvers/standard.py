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


# ------------------------------------------------------------------------------ module-level ops
# The reference's solver functions with its argument lists (standard.py:14, :85, :148, :168), for
# callers that use the module as an operator library.  Each runs on a native context built for
# the grid behind nconn (cached per configuration): the element sizes come from the element
# matrices, the constants the caller passes are checked against the context's.
_op_cache = {}


def _sizes_from_Ge(Ge):
    """(hx, hy) of a rectangular quad4 element: Ge[x][0][0] = -hy / 6, Ge[y][0][0] = -hx / 6."""
    G = np.asarray(nat.to_numpy(Ge))
    return float(-6.0 * G[1, 0, 0]), float(-6.0 * G[0, 0, 0])


def _mg_levels_for(nex, M_op):
    """Multigrid levels from the final mesh passed in place of the reference's M_op closure."""
    if M_op is None:
        raise ValueError("multigrid needs the final multigrid mesh [nex_f, ney_f] as M_op")
    return int(np.log2(nex // int(np.atleast_1d(M_op)[0]))) + 1


def _op_context(nconn, nn, hx, hy, *, gamma_d=None, gamma_p=None, set_zeroP=False, precond=None, M_op=None,
                tol=1e-8, bconn=None, nvec=None, gamma_d_phi=None, passive=False):
    from .._context import infer_grid
    nex, ney, periodic = infer_grid(nconn, int(nn))
    nlev = _mg_levels_for(nex, M_op) if precond == "multigrid" else 1
    gd = None if gamma_d is None else np.asarray(nat.to_numpy(gamma_d), dtype=np.int64)
    gp = None if gamma_p is None else np.atleast_1d(np.asarray(nat.to_numpy(gamma_p), dtype=np.int64))
    gdp = None if gamma_d_phi is None else np.asarray(nat.to_numpy(gamma_d_phi), dtype=np.int64)
    mg_pins = None
    if precond == "multigrid" and set_zeroP:
        if gp is None or list(gp) != [0]:
            raise NotImplementedError("a pinned multigrid from these arguments pins node 0 (the hierarchy's outlet "
                                      "pins of other meshes come from setup_solver)")
        mg_pins = [np.array([0], dtype=np.int64)] * nlev
    bc = None if bconn is None else np.asarray(nat.to_numpy(bconn), dtype=np.int64)
    nv = None if nvec is None else np.asarray(nat.to_numpy(nvec))
    key = (nex, ney, periodic, round(hx, 14), round(hy, 14), None if gd is None else gd.tobytes(),
           None if gp is None else gp.tobytes(), bool(set_zeroP), precond, nlev, float(tol),
           None if bc is None else bc.tobytes(), None if gdp is None else gdp.tobytes(), bool(passive))
    ctx = _op_cache.get(key)
    if ctx is None:
        ctx = Context(nex, ney, hx * nex, hy * ney, periodic, tol=tol, set_zeroP=set_zeroP, precond=precond,
                      mg_levels=nlev, gamma_d=gd, gamma_p=gp if (set_zeroP or precond == "approx_inv") else None,
                      mg_gamma_p=mg_pins, bconn=bc, nvec=nv, gamma_d_phi=gdp, approx_inv_variant=1 if passive else 0)
        _op_cache[key] = ctx
    return ctx


def _check_const(ctx, which, A):
    if A is not None and not np.allclose(ctx.element_constant(which), np.asarray(nat.to_numpy(A)), rtol=0,
                                         atol=1e-13 * max(1.0, float(np.abs(nat.to_numpy(A)).max()))):
        raise ValueError(f"{which} does not match the quad4 element of this grid")


def _theta_gamma(theta, gamma):
    if theta != 0.5 or gamma != 1:
        raise NotImplementedError("the native momentum solve is the four-step form: theta = 1/2, gamma = 1")


def solve_momentum(U_n, P_n, Me, Le, Ge, Ce_all_n, rho, mu, dt, nconn, nn, ndim, gamma_d, U_d, F, tol, theta,
                   gamma):
    """standard.py:14-43 (theta = 1/2, gamma = 1): BiCGSTAB on (rho/dt) M_c + theta (rho Ce + mu Le)
    with the Dirichlet rows replaced by the identity; returns (U_hat, iters)."""
    if ndim != 2:
        raise NotImplementedError("module-level solve_momentum is built for 2D quad4 grids")
    _theta_gamma(theta, gamma)
    hx, hy = _sizes_from_Ge(Ge)
    ctx = _op_context(nconn, nn, hx, hy, gamma_d=gamma_d, tol=tol)
    _check_const(ctx, "Me", Me)
    _check_const(ctx, "Le", Le)
    Ud = None if gamma_d is None else np.asarray(nat.to_numpy(U_d), dtype=np.float64)
    return ctx.solve_momentum(U_n, P_n, rho, mu, dt, U_d=Ud, F=F, Ce=Ce_all_n)


def solve_pressure(U_star, P_n, Le, Ge, rho, dt, nconn, nn, ndim, gamma_p, F_p, set_zeroP, precond, M_op, tol,
                   gamma):
    """standard.py:85-125 (gamma = 0): PCG on L with the gamma_p rows zeroed (set_zeroP) or the
    mean projected out; precond None / "jacobi" / "approx_inv" / "multigrid" (then M_op is the
    final multigrid mesh [nex_f, ney_f] in place of the reference's closure).  Returns (P, iters)."""
    if ndim != 2:
        raise NotImplementedError("module-level solve_pressure is built for 2D quad4 grids")
    if gamma != 0:
        raise NotImplementedError("the native pressure solve is the four-step form (gamma = 0)")
    hx, hy = _sizes_from_Ge(Ge)
    ctx = _op_context(nconn, nn, hx, hy, gamma_p=gamma_p, set_zeroP=set_zeroP, precond=precond, M_op=M_op,
                      tol=tol)
    _check_const(ctx, "Le", Le)
    return ctx.solve_pressure(U_star, P_n, F_p, rho, dt)


def four_step_timestep(U_n, P_n, M, Me, Le, Ge, rho, mu, dt, nconn, nn, ndim, wq, N, Nderiv, gamma_d, U_d, gamma_p,
                       F_p, F, set_zeroP, p_precond, p_M_op, tol, theta):
    """standard.py:148-165 (theta = 1/2); p_M_op as in solve_pressure.  Returns
    (U, P, U_iters, P_iters)."""
    if ndim != 2:
        raise NotImplementedError("module-level four_step_timestep is built for 2D quad4 grids")
    _theta_gamma(theta, 1)
    hx, hy = _sizes_from_Ge(Ge)
    ctx = _op_context(nconn, nn, hx, hy, gamma_d=gamma_d, gamma_p=gamma_p, set_zeroP=set_zeroP, precond=p_precond,
                      M_op=p_M_op, tol=tol)
    _check_const(ctx, "Me", Me)
    _check_const(ctx, "Le", Le)
    Ud = None
    if gamma_d is not None:
        Ud = nat.as_dev(np.asarray(nat.to_numpy(U_d), dtype=np.float64))
    return ctx.four_step_timestep(U_n, P_n, dt, rho, mu, Ud, F_p, F)


def compute(node_coords, bconn, nvec, boundary_shapefunc, gamma_d, U_d, nn, ndim, S_func, Me, nconn, t_n, t):
    """standard.py:167-188: F_p = int_Gamma_d q (U_d . n) and F = M_c (S(t_n) + S(t)) / 2."""
    if ndim != 2:
        raise NotImplementedError("module-level compute is built for 2D quad4 grids")
    X = np.asarray(nat.to_numpy(node_coords))
    nc = np.asarray(nat.to_numpy(nconn))
    hx, hy = float(X[nc[0, 1], 0] - X[nc[0, 0], 0]), float(X[nc[0, 3], 1] - X[nc[0, 0], 1])
    ctx = _op_context(nconn, nn, hx, hy, gamma_d=gamma_d, bconn=bconn, nvec=nvec)
    F_p, F = 0, 0
    if gamma_d is not None:
        F_p = ctx.boundary_integral(U_d)
    if S_func is not None:
        S = (nat.as_dev(S_func(t_n)) + nat.as_dev(S_func(t))) / 2
        Mek = ctx.element_constant("Me") if Me is None else np.asarray(nat.to_numpy(Me))
        F = ctx._empty(2 * int(nn))
        for d in range(2):
            ctx.elem_mult_const(Mek, S[d * nn:(d + 1) * nn], out=F[d * nn:(d + 1) * nn])
    return F_p, F
