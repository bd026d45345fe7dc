"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.

Tolerances (FP64 form): integer / index work bit-exact; single element and
matvec kernels 1e-13 relative (different summation order than the oracle's
BLAS); solves and timesteps 1e-10 relative L2, the accuracy set for the FP64
form, with identical iteration counts.
"""
import numpy as np
import pytest

from oracle import elements as oel
from oracle import emult as oem
from oracle import mesh as oms
from oracle import multigrid as omg
from oracle import ns

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(20241203)


def _np(x):
    return x.detach().cpu().numpy()


def _rel(a, b):
    return ns.rel_l2(a, b)


def _ctx(nex, ney, lx, ly, periodic, **kw):
    from jax_fem_cfd._context import Context
    print(f"\n[case] grid {nex}x{ney} periodic={periodic} {kw.get('precond')}", flush=True)
    return Context(nex, ney, lx, ly, periodic, **kw)


# ----------------------------------------------------------------------------------------- mesh
@pytest.mark.parametrize("nex,ney,periodic", [(7, 5, False), (16, 16, False), (8, 6, True), (64, 64, True)])
def test_device_mesh_bit_exact(nex, ney, periodic):
    c = _ctx(nex, ney, 2.0, 1.5, periodic)
    coords, nconn = c.mesh()
    f = oms.create_periodic_mesh_2d if periodic else oms.create_mesh_2d
    oc, on, *_ = f(nex, ney, 2.0, 1.5)
    assert np.array_equal(_np(nconn), on)
    assert np.array_equal(_np(coords), oc)


def test_element_constants_and_node_sums():
    c = _ctx(8, 8, 1.0, 1.0, False)
    coords, nconn, *_ = oms.create_mesh_2d(8, 8, 1.0, 1.0)
    wq, N, Nd = oel.quad4_shape_functions(coords[nconn[0]])
    assert np.allclose(c.element_constant("Me"), oel.mass_element(wq, N), rtol=0, atol=1e-17)
    assert np.allclose(c.element_constant("Le"), oel.laplacian_element_calc(wq, Nd), rtol=0, atol=1e-15)
    Ge = oel.gradient_element_calc(wq, N, Nd, 2)
    assert np.allclose(c.element_constant("Gx"), Ge[0], rtol=0, atol=1e-16)
    assert np.allclose(c.element_constant("Gy"), Ge[1], rtol=0, atol=1e-16)
    M, _ = oel.lumped_mass_matrix(nconn, wq, N, 81, 2)
    assert np.allclose(_np(c.lumped_mass()), M, rtol=1e-15, atol=0)
    Ld = oel.laplacian_diag(oel.laplacian_element_calc(wq, Nd), nconn, 81)
    assert np.allclose(_np(c.laplacian_diag()), Ld, rtol=1e-15, atol=0)


# ----------------------------------------------------------------------------------------- Ce
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("beta", [0.0, 0.5, 1.0])
def test_convection_matrices(periodic, beta):
    n = 12
    c = _ctx(n, n + 2, 1.0, 1.3, periodic)
    f = oms.create_periodic_mesh_2d if periodic else oms.create_mesh_2d
    coords, nconn, *_ = f(n, n + 2, 1.0, 1.3)
    nn = coords.shape[0]
    U = RNG.standard_normal(2 * nn)
    wq, N, Nd = oel.quad4_shape_functions(coords[nconn[0]])
    ref = oel.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, beta)
    got = _np(c.conv_elem_calc(U, beta))
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


# ----------------------------------------------------------------------------------------- matvecs
@pytest.mark.parametrize("periodic", [False, True])
def test_element_mult_api(periodic):
    from jax_fem_cfd.core import element_mult as em
    n = 10
    f = oms.create_periodic_mesh_2d if periodic else oms.create_mesh_2d
    coords, nconn, *_ = f(n, n + 3, 1.0, 1.3)
    nn = coords.shape[0]
    print(f"\n[case] element_mult periodic={periodic}", flush=True)
    wq, N, Nd = oel.quad4_shape_functions(coords[nconn[0]])
    Me, Le, Ge = oel.mass_element(wq, N), oel.laplacian_element_calc(wq, Nd), oel.gradient_element_calc(wq, N, Nd, 2)
    U, P = RNG.standard_normal(2 * nn), RNG.standard_normal(nn)
    Ce = oel.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, 0.5)
    Se = RNG.standard_normal(Ce.shape)
    tol = 1e-13
    assert _rel(_np(em.consistent_mass_mult(U, Me, nconn, nn, 2)), oem.consistent_mass_mult(U, Me, nconn, nn, 2)) < tol
    assert _rel(_np(em.laplacian_element_mult(P, Le, nconn, nn)), oem.laplacian_element_mult(P, Le, nconn, nn)) < tol
    Leoff = Le - np.diag(np.diag(Le))
    assert _rel(_np(em.laplacian_element_mult(P, Leoff, nconn, nn)), oem.laplacian_element_mult(P, Leoff, nconn, nn)) < tol
    assert _rel(_np(em.gradient_element_mult(P, Ge, nconn, nn, 2)), oem.gradient_element_mult(P, Ge, nconn, nn, 2)) < tol
    assert _rel(_np(em.divergence_element_mult(U, Ge, nconn, nn, 2)), oem.divergence_element_mult(U, Ge, nconn, nn, 2)) < tol
    assert _rel(_np(em.visc_conv_mult(U, Ce, Le, 0.01, 1.3, nconn, nn, 2)),
                oem.visc_conv_mult(U, Ce, Le, 0.01, 1.3, nconn, nn, 2)) < tol
    assert _rel(_np(em.scalar_visc_conv_mult(P, Ce, Le, 1e-3, nconn, nn)),
                oem.scalar_visc_conv_mult(P, Ce, Le, 1e-3, nconn, nn)) < tol
    assert _rel(_np(em.scalar_stabilized_visc_conv_mult(P, Ce, Se, Le, 1e-3, nconn, nn)),
                oem.scalar_stabilized_visc_conv_mult(P, Ce, Se, Le, 1e-3, nconn, nn)) < tol
    assert _rel(_np(em.scalar_consistent_mass_SUPG_mult(P, Se, Me, nconn, nn)),
                oem.scalar_consistent_mass_SUPG_mult(P, Se, Me, nconn, nn)) < tol


def test_element_calc_api():
    from jax_fem_cfd.core import element_calc as ec
    from jax_fem_cfd.core import shape_functions as sf
    coords, nconn, *_ = oms.create_mesh_2d(9, 7, 0.9, 0.7)
    nn = coords.shape[0]
    wq, N, Nd = sf.quad4_shape_functions(coords[nconn[0]])
    U = RNG.standard_normal(2 * nn)
    got = _np(ec.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, 0.5))
    ref = oel.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, 0.5)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    M, Me = ec.lumped_mass_matrix(nconn, wq, N, nn, 2)
    Mo, _ = oel.lumped_mass_matrix(nconn, wq, N, nn, 2)
    assert np.allclose(_np(M), Mo, rtol=1e-15, atol=0)
    assert np.allclose(_np(ec.laplacian_diag(ec.laplacian_element_calc(wq, Nd), nconn, nn)),
                       oel.laplacian_diag(oel.laplacian_element_calc(wq, Nd), nconn, nn), rtol=1e-15, atol=0)


def test_boundary_integral():
    from jax_fem_cfd.core import element_calc as ec
    s, surf, h, U_d = ns.build([8, 6], [1.0, 0.75], False, [0, 1, 3], [], 1e-8, True, None, None)
    U_d = U_d + RNG.standard_normal(U_d.shape)
    ref = oel.u_dot_n_boundary_integral_1d(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords)
    got = _np(ec.u_dot_n_boundary_integral_1d(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords))
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


# ----------------------------------------------------------------------------------------- multigrid
@pytest.mark.parametrize("periodic,set_zeroP", [(False, True), (False, False), (True, False)])
def test_vcycle(periodic, set_zeroP):
    n, nf = 32, 4
    s, *_ = ns.build([n, n], [1.0, 1.0], periodic, [], [], 1e-10, set_zeroP, "multigrid", [nf, nf])
    c = _ctx(n, n, 1.0, 1.0, periodic, tol=1e-10, set_zeroP=set_zeroP, precond="multigrid",
             mg_levels=len(s.levels), gamma_p=[0] if set_zeroP else None, mg_gamma_p=[L["gamma_p"] for L in s.levels])
    r = RNG.standard_normal(s.nn)
    if set_zeroP:
        r[0] = 0
    else:
        r -= r.mean()
    ref = s.M_op(r)
    got = _np(c.vcycle(r))
    assert _rel(got, ref) < 1e-12


@pytest.mark.parametrize("set_zeroP,nf", [(True, 4), (True, 16), (False, 8)])
def test_vcycle_mixed_form(set_zeroP, nf):
    """FP32 V-cycle of the mixed form against the oracle's FP64 V-cycle.  Pinned: the coarse
    solve is the exact (dense, FP32) inverse of the pinned coarse operator, the oracle's is CG to
    1e-12 -- both solve the coarse system, so only FP32 rounding separates them (2e-5)."""
    n = 64
    s, *_ = ns.build([n, n], [1.0, 1.0], False, [], [], 1e-12, set_zeroP, "multigrid", [nf, nf])
    c = _ctx(n, n, 1.0, 1.0, False, tol=1e-12, set_zeroP=set_zeroP, precond="multigrid", mg_levels=len(s.levels),
             precision="mixed", gamma_p=[0] if set_zeroP else None,
             mg_gamma_p=[L["gamma_p"] for L in s.levels])
    r = RNG.standard_normal(s.nn)
    if set_zeroP:
        r[0] = 0
    else:
        r -= r.mean()
    ref = s.M_op(r)
    got = _np(c.vcycle(r))
    assert _rel(got, ref) < 2e-5, _rel(got, ref)


def test_vcycle_api_tuple():
    from jax_fem_cfd.core import multigrid as mg
    from jax_fem_cfd.core import shape_functions as sf
    from jax_fem_cfd.setup import regular_mesh as rm
    data = mg.multigrid_setup([32, 32], [4, 4], [1, 1], [], rm.create_mesh_2d, sf.quad4_shape_functions, 1, None)
    s, *_ = ns.build([32, 32], [1.0, 1.0], False, [], [], 1e-8, True, "multigrid", [4, 4])
    r = RNG.standard_normal(s.nn)
    r[0] = 0
    got = _np(mg.v_cycle(r, *data, 1e-8, False, True))
    assert _rel(got, s.M_op(r)) < 1e-10


# ----------------------------------------------------------------------------------------- solvers
@pytest.mark.parametrize("set_zeroP", [True, False])
def test_pressure_cg_multigrid(set_zeroP):
    n = 64
    s, *_ = ns.build([n, n], [1.0, 1.0], False, [], [], 1e-10, set_zeroP, "multigrid", [8, 8])
    c = _ctx(n, n, 1.0, 1.0, False, tol=1e-10, set_zeroP=set_zeroP, precond="multigrid", mg_levels=len(s.levels),
             gamma_p=[0] if set_zeroP else None, mg_gamma_p=[L["gamma_p"] for L in s.levels])
    b = RNG.standard_normal(s.nn)
    if set_zeroP:
        b[0] = 0
    else:
        b -= b.mean()

    def A(x):
        y = oem.laplacian_element_mult(x, s.Le, s.nconn, s.nn)
        if set_zeroP:
            y[0] = 0
        return y

    from oracle.solvers import cg
    xr, kr = cg(A, b, tol=1e-10, maxiter=1000, M=s.M_op)
    x, k, conv = c.pressure_cg(b, tol=1e-10, maxiter=1000)
    assert conv and k == kr
    assert _rel(_np(x), xr) < 1e-10


def test_neumann_pressure_cg_mixed_keeps_the_mean_of_x0():
    """Pure Neumann, mixed form (the tiled V-cycle without the intermediate mean projections,
    whose z carries a constant): the solve still returns the reference's iterate, which keeps the
    mean of x0 -- x0 with a nonzero mean, the raw solution against the oracle's CG from the same
    x0."""
    n = 64
    s, *_ = ns.build([n, n], [1.0, 1.0], False, [], [], 1e-10, False, "multigrid", [8, 8])
    c = _ctx(n, n, 1.0, 1.0, False, tol=1e-10, set_zeroP=False, precond="multigrid", mg_levels=len(s.levels),
             precision="mixed", mg_gamma_p=[L["gamma_p"] for L in s.levels])
    b = RNG.standard_normal(s.nn)
    b -= b.mean()
    x0 = RNG.standard_normal(s.nn) + 0.75

    def A(x):
        return oem.laplacian_element_mult(x, s.Le, s.nconn, s.nn)

    from oracle.solvers import cg
    xr, kr = cg(A, b, x0=x0.copy(), tol=1e-10, maxiter=1000, M=s.M_op)
    x, k, conv = c.pressure_cg(b, x0=x0, tol=1e-10, maxiter=1000)
    xg = _np(x)
    assert conv
    assert abs(xg.mean() - x0.mean()) < 1e-12 * abs(x0.mean())
    assert _rel(xg, xr) < 1e-8, _rel(xg, xr)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_vcycle_and_pressure_cg_at_512(precision):
    """512x512 elements (263,169 nodes), pinned: the sizes at which the V-cycle runs its staged
    fine-level kernels (the fused down leg with 4 coarse rows per block from 200 coarse rows on,
    the LDS-staged up-leg Jacobi from 2^17 nodes on) and the PCG its staged Laplacian.  FP64: the
    V-cycle to 1e-12 and the PCG with equal iterations to 1e-10 against the oracle.  Mixed: the
    FP32 V-cycle to 2e-5, and the flexible PCG converged to the same tolerance (solution 1e-7)."""
    n, nf = 512, 16
    s, *_ = ns.build([n, n], [1.0, 1.0], False, [], [], 1e-10, True, "multigrid", [nf, nf])
    print(f"\n[case] V-cycle and PCG {n}x{n} {precision}", flush=True)
    c = _ctx(n, n, 1.0, 1.0, False, tol=1e-10, set_zeroP=True, precond="multigrid", mg_levels=len(s.levels),
             precision=precision, gamma_p=[0], mg_gamma_p=[L["gamma_p"] for L in s.levels])
    r = RNG.standard_normal(s.nn)
    r[0] = 0
    got = _np(c.vcycle(r))
    assert _rel(got, s.M_op(r)) < (1e-12 if precision == "fp64" else 2e-5)

    def A(x):
        y = oem.laplacian_element_mult(x, s.Le, s.nconn, s.nn)
        y[0] = 0
        return y

    from oracle.solvers import cg
    xr, kr = cg(A, r, tol=1e-10, maxiter=1000, M=s.M_op)
    x, k, conv = c.pressure_cg(r, tol=1e-10, maxiter=1000)
    print(f"  iterations {k} oracle {kr}  rel x {_rel(_np(x), xr):.2e}", flush=True)
    assert conv
    if precision == "fp64":
        assert k == kr and _rel(_np(x), xr) < 1e-10
    else:
        assert _rel(_np(x), xr) < 1e-7


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_pressure_cg_approx_inv(set_zeroP):
    """PCG with the two-term approximate inverse (standard.py:225-234); its gamma_p rows (node 0
    without outlets) are zeroed in both the pinned and the pure-Neumann form, as the reference's."""
    n = 48
    s, *_ = ns.build([n, n], [1.0, 1.0], False, [], [], 1e-10, set_zeroP, "approx_inv", None)
    c = _ctx(n, n, 1.0, 1.0, False, tol=1e-10, set_zeroP=set_zeroP, precond="approx_inv", gamma_p=[0])
    b = RNG.standard_normal(s.nn)
    if set_zeroP:
        b[0] = 0
    else:
        b -= b.mean()

    def A(x):
        y = oem.laplacian_element_mult(x, s.Le, s.nconn, s.nn)
        if set_zeroP:
            y[0] = 0
        return y

    from oracle.solvers import cg
    xr, kr = cg(A, b, tol=1e-10, maxiter=1000, M=s.M_op)
    x, k, conv = c.pressure_cg(b, tol=1e-10, maxiter=1000)
    assert k == kr and conv == (kr < 1000)
    assert _rel(_np(x), xr) < 1e-10


def test_cavity_approx_inv_steps_match_oracle():
    """The full step with the approx_inv pressure preconditioner (cavity 32x32, 3 steps)."""
    import jax_fem_cfd.setup.simulation as sim
    n = 32
    cfg = sim.Config(mesh="simple", solver="standard", mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                     solver_config=sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, 1e-10, True, "approx_inv",
                                                            [4, 4]))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    U_d = sim.set_face_velocity(U_d, surf[2][1:-1], gamma_d, 1.0, nn, 1)
    s, osurf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, "approx_inv", None)
    oUd = ns.cavity_lid(s, osurf, oUd)
    U = np.zeros(2 * nn)
    U[gamma_d] = U_d
    P = np.zeros(nn)
    Uo, Po = U.copy(), P.copy()
    for _ in range(3):
        F_p, F = compute(U_d, None, 0, 0)
        U, P, ui, pi = timestep(U, P, 2e-3, 1.0, 1e-2, U_d, F_p, F)
        Uo, Po, uio, pio = ns.four_step_timestep(s, Uo, Po, 2e-3, 1.0, 1e-2, oUd, 0, 0)
        assert (ui, pi) == (uio, pio)
    assert _rel(_np(U), Uo) < 1e-10 and _rel(_np(P), Po) < 1e-10


def test_generic_solver_wrappers():
    from jax_fem_cfd.utils import jax_iterative_solvers as jis
    from oracle import solvers as osv
    import torch
    n = 50
    A = np.diag(np.linspace(1, 10, n)) + 0.1 * RNG.standard_normal((n, n))
    S = A + A.T + 20 * np.eye(n)
    b = RNG.standard_normal(n)
    At, St = torch.tensor(A, device="cuda"), torch.tensor(S, device="cuda")
    x, k = jis.cg(lambda v: St @ v, b, tol=1e-10)
    xr, kr = osv.cg(lambda v: S @ v, b, tol=1e-10)
    assert k == kr and _rel(_np(x), xr) < 1e-9
    x, k = jis.bicgstab(lambda v: At @ v, b, tol=1e-10)
    xr, kr = osv.bicgstab(lambda v: A @ v, b, tol=1e-10)
    assert k == kr and _rel(_np(x), xr) < 1e-8


def _cavity(n, tol, precision="fp64", ce_format=None, mu=1e-3):
    import jax_fem_cfd.setup.simulation as sim
    cfg = sim.Config(mesh="simple", solver="standard",
                     mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                     solver_config=sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, tol, True, "multigrid",
                                                            [4, 4], precision=precision, ce_format=ce_format))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    U_d = sim.set_face_velocity(U_d, surf[2][1:-1], gamma_d, 1.0, nn, 1)
    U0 = np.random.default_rng(1).normal(0, 1e-3, 2 * nn)
    U0[gamma_d] = U_d
    return compute, timestep, U_d, U0, np.zeros(nn), nn


def test_momentum_solve_matches_oracle():
    n = 32
    compute, timestep, U_d, U0, P0, nn = _cavity(n, 1e-10)
    ctx = timestep.context
    s, surf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, "multigrid", [4, 4])
    oUd = ns.cavity_lid(s, surf, oUd)
    P0 = RNG.standard_normal(nn) * 0.1
    Ce = oel.conservative_convec_elem_calc(s.nconn, U0, s.wq, s.N, s.Nderiv, nn, 2, 0.5)
    ref, kr = ns.solve_momentum(U0, P0, s.Me, s.Le, s.Ge, Ce, 1.0, 1e-3, 5e-4, s.nconn, nn, 2, s.gamma_d, oUd, 0,
                                1e-10, 0.5, 1)
    got, k = ctx.solve_momentum(U0, P0, 1.0, 1e-3, 5e-4, U_d=U_d)
    assert k == kr and _rel(_np(got), ref) < 1e-10
    got2, k2 = ctx.solve_momentum(U0, P0, 1.0, 1e-3, 5e-4, U_d=U_d, Ce=Ce)
    assert k2 == kr and _rel(_np(got2), ref) < 1e-10


# ----------------------------------------------------------------------------------------- timestep
def test_taylor_green_64_100_steps():
    """Config 1: 2D periodic Taylor-Green, 64x64 Q1, nu = 0.01, dt = 1e-3, 100 steps,
    FP64, MG-PCG and BiCGSTAB tol 1e-10 -- per-field error vs the oracle."""
    import jax_fem_cfd.setup.simulation as sim
    rho, mu, dt, steps = 1.0, 0.01, 1e-3, 100
    cfg = sim.Config(mesh="simple", solver="standard",
                     mesh_config=sim.SimpleMeshConfig([64, 64], [2 * np.pi, 2 * np.pi], 1, True),
                     solver_config=sim.StandardSolverConfig([], [], [], False, 1e-10, False, "multigrid", [2, 2]))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    u, v, p = ns.taylor_green(coords, 0, rho, mu)
    U0, P0 = np.concatenate([u, v]), p - p.mean()
    s, *_ = ns.build([64, 64], [2 * np.pi, 2 * np.pi], True, [], [], 1e-10, False, "multigrid", [2, 2])
    U, P = U0, P0
    Uo, Po = U0.copy(), P0.copy()
    its, itso = [], []
    for _ in range(steps):
        F_p, F = compute(U_d, None, 0, 0)
        U, P, ui, pi = timestep(U, P, dt, rho, mu, U_d, F_p, F)
        Uo, Po, uio, pio = ns.four_step_timestep(s, Uo, Po, dt, rho, mu, None, 0, 0)
        its.append((ui, pi))
        itso.append((uio, pio))
    Ug, Pg = _np(U), _np(P)
    eu = _rel(Ug[:nn], Uo[:nn])
    ev = _rel(Ug[nn:], Uo[nn:])
    ep = _rel(Pg, Po)
    print(f"TG64 100 steps: relL2 u {eu:.2e} v {ev:.2e} p {ep:.2e}; iters equal: {its == itso}")
    assert its == itso
    assert eu < 1e-10 and ev < 1e-10 and ep < 1e-10
    ue, ve, pe = ns.taylor_green(coords, steps * dt, rho, mu)
    assert _rel(Ug, np.concatenate([ue, ve])) < 1e-4


def test_cavity_steps_match_oracle_and_are_deterministic():
    n = 64
    compute, timestep, U_d, U0, P0, nn = _cavity(n, 1e-8)
    s, surf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-8, True, "multigrid", [4, 4])
    oUd = ns.cavity_lid(s, surf, oUd)
    U, P, Uo, Po = U0, P0, U0.copy(), P0.copy()
    for _ in range(5):
        F_p, F = compute(U_d, None, 0, 0)
        U, P, ui, pi = timestep(U, P, 5e-4, 1.0, 1e-3, U_d, F_p, F)
        Uo, Po, uio, pio = ns.four_step_timestep(s, Uo, Po, 5e-4, 1.0, 1e-3, oUd, 0, 0)
        assert (ui, pi) == (uio, pio)
    assert _rel(_np(U), Uo) < 1e-10 and _rel(_np(P), Po) < 1e-10
    # bitwise run-to-run reproducibility (no float atomics)
    Ua, Pa, *_ = timestep(U0, P0, 5e-4, 1.0, 1e-3, U_d, 0, 0)
    Ub, Pb, *_ = timestep(U0, P0, 5e-4, 1.0, 1e-3, U_d, 0, 0)
    assert np.array_equal(_np(Ua), _np(Ub)) and np.array_equal(_np(Pa), _np(Pb))


def test_step_whose_momentum_chunk_is_too_short():
    """A fresh context predicts one BiCGSTAB pair (2 iterations) for its first momentum solve;
    a first step from a strongly perturbed state at tol 1e-12 needs more, so the step settles
    its momentum solve after the pressure solve, completes it and redoes U* and the pressure
    solve.  Fields and iteration counts must be those of the oracle's step."""
    n = 64
    compute, timestep, U_d, U0, P0, nn = _cavity(n, 1e-12)
    s, surf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-12, True, "multigrid", [4, 4])
    oUd = ns.cavity_lid(s, surf, oUd)
    U0 = U0 + np.random.default_rng(3).normal(0, 0.1, U0.shape)
    Uo, Po, uio, pio = ns.four_step_timestep(s, U0.copy(), P0.copy(), 5e-4, 1.0, 1e-3, oUd, 0, 0)
    U, P, ui, pi = timestep(U0, P0, 5e-4, 1.0, 1e-3, U_d, 0, 0)
    print(f"\n[case] short momentum chunk: iters {ui},{pi} oracle {uio},{pio}", flush=True)
    assert uio > 2 and (ui, pi) == (uio, pio)
    assert _rel(_np(U), Uo) < 1e-10 and _rel(_np(P), Po) < 1e-10


# per-field relative L2 bounds of the mixed form against the FP64 oracle after 5 steps at tol 1e-8:
# (velocity, pressure): FP32 Ce within the solver tolerance; FP16 + scale (11-bit entries) and
# BF16 + scale (8-bit entries) at the rounding level of their stored convection matrices, whose
# pressure error grows by about 4e-6 (FP16) per step over the first steps from rest
def test_in_place_step_whose_momentum_chunk_is_too_short():
    """The same settled-and-redone step with U == U_n and P == P_n (one buffer pair updated in
    place): the redo must build U* and the pressure solve's x0 from P_n, so P may be written
    only after the momentum solve has settled.  Fields and counts equal the oracle's step."""
    import torch
    n = 64
    compute, timestep, U_d, U0, P0, nn = _cavity(n, 1e-12)
    s, surf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-12, True, "multigrid", [4, 4])
    oUd = ns.cavity_lid(s, surf, oUd)
    U0 = U0 + np.random.default_rng(3).normal(0, 0.1, U0.shape)
    P0 = np.random.default_rng(4).normal(0, 0.1, P0.shape)
    Uo, Po, uio, pio = ns.four_step_timestep(s, U0.copy(), P0.copy(), 5e-4, 1.0, 1e-3, oUd, 0, 0)
    ctx = timestep.context
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    P = torch.tensor(P0, dtype=torch.float64, device="cuda")
    Ud = torch.tensor(U_d, dtype=torch.float64, device="cuda")
    U2, P2, ui, pi = ctx.four_step_timestep(U, P, 5e-4, 1.0, 1e-3, Ud, out=(U, P))
    torch.cuda.synchronize()
    print(f"\n[case] in-place short momentum chunk: iters {ui},{pi} oracle {uio},{pio}", flush=True)
    assert U2.data_ptr() == U.data_ptr() and P2.data_ptr() == P.data_ptr()
    assert uio > 2 and (ui, pi) == (uio, pio)
    assert _rel(_np(U), Uo) < 1e-10 and _rel(_np(P), Po) < 1e-10


MIXED_TOL = {"f32": (1e-8, 1e-8), "f16s": (2e-5, 1e-4), "bf16s": (2e-4, 1e-3)}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_four_step_rollout_equals_single_steps(precision):
    """The native step loop (fcfd_four_step_rollout) runs the same kernels as one
    fcfd_four_step_timestep call per step: after 5 steps the fields are bitwise equal and the
    per-step iteration counts identical."""
    n = 64
    c, t, U_d, U0, P0, nn = _cavity(n, 1e-8, precision=precision)
    print(f"\n[case] rollout {n} {precision}", flush=True)
    U, P, its = U0, P0, []
    for _ in range(5):
        U, P, ui, pi = t(U, P, 5e-4, 1.0, 1e-3, U_d, 0, 0)
        its.append((ui, pi))
    Ur, Pr, itr = t.context.four_step_rollout(U0, P0, 5, 5e-4, 1.0, 1e-3, U_d, None, None)
    assert np.array_equal(_np(Ur), _np(U)) and np.array_equal(_np(Pr), _np(P))
    assert [tuple(map(int, r)) for r in itr] == its


def test_rollout_with_graphs_equals_direct_launches(monkeypatch):
    """FCFD_GRAPHS=1 captures the solver chunks into HIP graphs and replays them; the fields
    (bitwise) and per-step counts equal those of the default direct launches."""
    n = 64
    c, t, U_d, U0, P0, nn = _cavity(n, 1e-8, precision="mixed")
    Ur, Pr, itr = t.context.four_step_rollout(U0, P0, 6, 5e-4, 1.0, 1e-3, U_d, None, None)
    monkeypatch.setenv("FCFD_GRAPHS", "1")
    c1, t1, *_ = _cavity(n, 1e-8, precision="mixed")
    monkeypatch.delenv("FCFD_GRAPHS")
    print(f"\n[case] rollout {n} mixed with captured graphs", flush=True)
    Ug, Pg, itg = t1.context.four_step_rollout(U0, P0, 6, 5e-4, 1.0, 1e-3, U_d, None, None)
    assert [tuple(map(int, r)) for r in itg] == [tuple(map(int, r)) for r in itr]
    assert np.array_equal(_np(Ug), _np(Ur)) and np.array_equal(_np(Pg), _np(Pr))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rollout_whose_chunks_all_fall_short(precision, monkeypatch):
    """FCFD_SPEC_CHUNK=1 gives every solve of every speculative step a one-iteration chunk, so each
    step of the native loop falls short in its momentum or its pressure solve: the steps enqueued
    behind it do nothing, the step is resumed on the host and the next one enqueued again.  Fields
    (bitwise) and per-step counts must equal those of single steps on a context with predicted
    chunks, and the counts must exceed one iteration in both solves (both fallbacks taken)."""
    n = 64
    c, t, U_d, U0, P0, nn = _cavity(n, 1e-10, precision=precision)
    U, P, its = U0, P0, []
    for _ in range(6):
        U, P, ui, pi = t(U, P, 5e-4, 1.0, 1e-3, U_d, 0, 0)
        its.append((ui, pi))
    monkeypatch.setenv("FCFD_SPEC_CHUNK", "1")
    c1, t1, *_ = _cavity(n, 1e-10, precision=precision)
    monkeypatch.delenv("FCFD_SPEC_CHUNK")
    print(f"\n[case] rollout with one-iteration chunks {n} {precision}: counts {its}", flush=True)
    Ur, Pr, itr = t1.context.four_step_rollout(U0, P0, 6, 5e-4, 1.0, 1e-3, U_d, None, None)
    assert [tuple(map(int, r)) for r in itr] == its
    assert np.array_equal(_np(Ur), _np(U)) and np.array_equal(_np(Pr), _np(P))
    assert any(ui > 1 for ui, _ in its) and all(pi > 1 for _, pi in its)


@pytest.mark.parametrize("ce_format", ["f32", "f16s", "bf16s"])
def test_mixed_precision_cavity(ce_format):
    """Mixed form (FP32 V-cycle in FP64 flexible CG, reduced-precision stored Ce, FP64 Krylov
    vectors, dots and projection): after each of 5 steps at tol 1e-8 the velocity components (u, v)
    and the pressure stay within MIXED_TOL[ce_format] relative L2 of the FP64 oracle's step."""
    n = 64
    cm, tm, U_d, U0, P0, nn = _cavity(n, 1e-8, precision="mixed", ce_format=ce_format)
    s, surf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-8, True, "multigrid", [4, 4])
    oUd = ns.cavity_lid(s, surf, oUd)
    print(f"\n[case] mixed cavity {n} Ce {ce_format}", flush=True)
    Uo, Po, Um, Pm = U0.copy(), P0.copy(), U0, P0
    for k in range(5):
        Uo, Po, uio, pio = ns.four_step_timestep(s, Uo, Po, 5e-4, 1.0, 1e-3, oUd, 0, 0)
        Um, Pm, ui, pi = tm(Um, Pm, 5e-4, 1.0, 1e-3, U_d, 0, 0)
        Ug, Pg = _np(Um), _np(Pm)
        err = {"u": _rel(Ug[:nn], Uo[:nn]), "v": _rel(Ug[nn:], Uo[nn:]), "p": _rel(Pg, Po)}
        print(f"  step {k + 1}: iters ({ui},{pi}) oracle ({uio},{pio})  " +
              "  ".join(f"{f} {e:.2e}" for f, e in err.items()), flush=True)
        tu, tp = MIXED_TOL[ce_format]
        assert err["u"] < tu and err["v"] < tu and err["p"] < tp


# ----------------------------------------------------------------------------------------- passive scalar
def _rect_phi(coords, x0=0.3, x1=0.7, y0=0.2, y1=0.6):
    x, y = coords[:, 0], coords[:, 1]
    return ((x >= x0) & (x <= x1) & (y >= y0) & (y <= y1)).astype(float)


@pytest.mark.parametrize("supg", [False, True])
def test_passive_transport_matches_oracle(supg):
    """Passive scalar (passive_transport.py:64-97): cavity 32x32 with a swirl, scalar
    rectangle, 4 steps, plain and SUPG -- U, P, phi vs the oracle with equal iteration counts."""
    import jax_fem_cfd.setup.simulation as sim
    n, dt, rho, mu, D = 32, 2e-3, 1.0, 1e-2, 1e-4
    cfg = sim.Config(mesh="simple", solver="passive_transport",
                     mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                     solver_config=sim.PassiveTransportConfig(
                         sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, 1e-10, True, "multigrid", [4, 4]),
                         sim.ADEConfig([], supg)))
    compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi = sim.setup(cfg)
    print(f"\n[case] passive transport supg={supg}", flush=True)
    x, y = coords[:, 0], coords[:, 1]
    U0 = np.concatenate([np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y)])
    U0[gamma_d] = U_d
    phi0 = _rect_phi(coords)
    s, osurf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, "multigrid", [4, 4])
    U, P, phi = U0, np.zeros(nn), phi0
    Uo, Po, phio = U0.copy(), np.zeros(nn), phi0.copy()
    for _ in range(4):
        F_p, F = compute(U_d, None, 0, 0)
        U, P, phi, ui, pi, fi = timestep(U, P, phi, dt, rho, mu, D, U_d, None, F_p, F, None)
        Uo, Po, phio, uio, pio, fio = ns.passive_timestep(s, Uo, Po, phio, dt, rho, mu, D, oUd, None, 0, 0, None,
                                                          supg)
        assert (ui, pi, fi) == (uio, pio, fio)
    assert _rel(_np(U), Uo) < 1e-10 and _rel(_np(P), Po) < 1e-10
    assert _rel(_np(phi), phio) < 1e-10


def test_passive_transport_approx_inv_matches_oracle():
    """The passive solver with approx_inv: its own closure (passive_transport.py:154-162) keeps
    D1x / D1L on the gamma_p rows, unlike the standard solver's (standard.py:225-234)."""
    import jax_fem_cfd.setup.simulation as sim
    n, dt, rho, mu, D = 32, 2e-3, 1.0, 1e-2, 1e-4
    cfg = sim.Config(mesh="simple", solver="passive_transport",
                     mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                     solver_config=sim.PassiveTransportConfig(
                         sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, 1e-10, True, "approx_inv"),
                         sim.ADEConfig([], False)))
    compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi = sim.setup(cfg)
    print("\n[case] passive transport approx_inv", flush=True)
    x, y = coords[:, 0], coords[:, 1]
    U0 = np.concatenate([np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y)])
    U0[gamma_d] = U_d
    phi0 = _rect_phi(coords)
    s, osurf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, "approx_inv", None,
                                passive=True)
    U, P, phi = U0, np.zeros(nn), phi0
    Uo, Po, phio = U0.copy(), np.zeros(nn), phi0.copy()
    for _ in range(3):
        F_p, F = compute(U_d, None, 0, 0)
        U, P, phi, ui, pi, fi = timestep(U, P, phi, dt, rho, mu, D, U_d, None, F_p, F, None)
        Uo, Po, phio, uio, pio, fio = ns.passive_timestep(s, Uo, Po, phio, dt, rho, mu, D, oUd, None, 0, 0, None,
                                                          False)
        assert (ui, pi, fi) == (uio, pio, fio)
    assert _rel(_np(U), Uo) < 1e-10 and _rel(_np(P), Po) < 1e-10 and _rel(_np(phi), phio) < 1e-10


def test_supg_element_calc_api():
    from jax_fem_cfd.core import element_calc as ec
    from jax_fem_cfd.core import shape_functions as sf
    coords, nconn, *_ = oms.create_mesh_2d(8, 6, 1.0, 0.75)
    nn = coords.shape[0]
    wq, N, Nd = sf.quad4_shape_functions(coords[nconn[0]])
    U = RNG.standard_normal(2 * nn)
    Ct = _np(ec.Ctau_tilde_elem_calc(nconn, U, wq, N, Nd, nn, 2))
    Cto = oel.Ctau_tilde_elem_calc(nconn, U, wq, N, Nd, nn, 2)
    assert np.abs(Ct - Cto).max() <= 1e-13 * np.abs(Cto).max()
    Ce = oel.conservative_convec_elem_calc(nconn, U, wq, N, Nd, nn, 2, 1.0)
    tau = _np(ec.calculate_tau(nconn, U, Ce, Cto, N, nn, 2, 1e-4, 1e-3, 2))
    tauo = oel.calculate_tau(nconn, U, Ce, Cto, N, nn, 2, 1e-4, 1e-3, 2)
    assert _rel(tau, tauo) < 1e-13


# ----------------------------------------------------------------------------------------- host loop
def test_timestep_loop_matches_oracle_loop():
    """timestep_loop (simulation.py:229-360): step count rule, compute -> timestep per step,
    iteration counts and solutions recorded on save steps and the last step."""
    import jax_fem_cfd.setup.simulation as sim
    n, dt = 16, 5e-3
    cfg = sim.Config(mesh="simple", solver="standard", mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                     solver_config=sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, 1e-10, True, "multigrid",
                                                            [4, 4]))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    U_d = sim.set_face_velocity(U_d, surf[2][1:-1], gamma_d, 1.0, nn, 1)
    U0 = np.random.default_rng(3).normal(0, 1e-2, 2 * nn)
    U0[gamma_d] = U_d
    t_final = 12 * dt
    U, P, ui, pi, steps, U_list = sim.timestep_loop(t_final, dt, 5, 100, compute, timestep, [1.0, 1e-2], [U_d],
                                                    [U0, np.zeros(nn)])
    s, osurf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, "multigrid", [4, 4])
    oUd = ns.cavity_lid(s, osurf, oUd)
    oU, oP, oui, opi, osteps, oU_list = ns.timestep_loop(s, t_final, dt, 5, 1.0, 1e-2, oUd, U0, np.zeros(nn))
    assert np.array_equal(steps, osteps) and np.array_equal(ui, oui) and np.array_equal(pi, opi)
    assert U_list.shape == oU_list.shape
    assert _rel(U_list, oU_list) < 1e-10 and _rel(_np(P), oP) < 1e-10


def test_timestep_loop_passive_returns_reference_tuple():
    import jax_fem_cfd.setup.simulation as sim
    n, dt, rho, mu, D = 16, 5e-3, 1.0, 1e-2, 1e-3
    cfg = sim.Config(mesh="simple", solver="passive_transport",
                     mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                     solver_config=sim.PassiveTransportConfig(
                         sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, 1e-10, True, "multigrid", [4, 4]),
                         sim.ADEConfig([], False)))
    compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi = sim.setup(cfg)
    U_d = sim.set_face_velocity(U_d, surf[2][1:-1], gamma_d, 1.0, nn, 1)
    U0 = np.zeros(2 * nn)
    U0[gamma_d] = U_d
    phi0 = _rect_phi(coords)
    out = sim.timestep_loop(7 * dt, dt, 3, 100, compute, timestep, [rho, mu, D], [U_d, None],
                            [U0, np.zeros(nn), phi0])
    U, P, phi, ui, pi, fi, steps, U_list, phi_list = out
    assert list(steps) == [3, 6, 7] and len(fi) == 3 and phi_list.shape == (4, nn) and U_list.shape == (4, 2 * nn)
    s, osurf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, "multigrid", [4, 4])
    oUd = ns.cavity_lid(s, osurf, oUd)
    oU, oP, ophi = U0.copy(), np.zeros(nn), phi0.copy()
    for _ in range(7):
        oU, oP, ophi, *_ = ns.passive_timestep(s, oU, oP, ophi, dt, rho, mu, D, oUd, None, 0, 0, None, False)
    assert _rel(_np(phi), ophi) < 1e-10 and _rel(phi_list[-1], ophi) < 1e-10
