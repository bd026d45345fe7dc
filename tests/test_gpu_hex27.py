"""GPU parity of the triquadratic (hex27) path against the oracle: element constants, lumped and
consistent mass, convection matrices, the quad9-face boundary integral, PCG, and whole four-step
timesteps of a channel (parabolic inflow, outlet, no-slip walls) with each hex27 pressure
preconditioner.  The reference has no 3D quadratic element, so the oracle's hex27 pieces are
pinned by tests/test_oracle_hex27.py (closed forms, independent quadrature)."""
import numpy as np
import pytest

from oracle import elements as el
from oracle import emult
from oracle import mesh as ms
from oracle import ns

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(2728)


def _np(x):
    return x.detach().cpu().numpy()


def _rel(a, b):
    return ns.rel_l2(a, b)


def _channel(n=(4, 2, 2), L=(2.0, 1.0, 1.0), precond="jacobi", tol=1e-10, precision="fp64", final=None):
    import jax_fem_cfd.setup.simulation as sim
    cfg = sim.Config(mesh="simple", solver="standard",
                     mesh_config=sim.SimpleMeshConfig(list(n), list(L), 2, False),
                     solver_config=sim.StandardSolverConfig([0, 2, 3, 4, 5], [], [1], False, tol, True, precond,
                                                            final, precision))
    out = sim.setup(cfg)
    s, surf, _, _ = ns.build(list(n), list(L), False, [0, 2, 3, 4, 5], [1], tol, True, precond, final,
                             shape_func_ord=2)
    compute, timestep, coords, osurf, h, nn, gamma_d, U_d = out
    assert nn == s.nn and np.allclose(coords, s.coords)
    assert np.array_equal(np.asarray(gamma_d), s.gamma_d)
    U_d = ns.channel_inflow(s, surf, np.zeros(len(s.gamma_d)), L[1], L[2])
    return compute, timestep, s, surf, U_d


def test_hex27_constants_match_oracle():
    print("\n[case] hex27 element constants", flush=True)
    _, timestep, s, _, _ = _channel((3, 2, 2), (1.5, 1.0, 0.5))
    ctx = timestep.context
    assert ctx.ndim == 3 and ctx.nen == 27 and ctx.nn == s.nn
    assert np.allclose(ctx.element_constant("Me"), s.Me, rtol=0, atol=1e-15 * np.abs(s.Me).max())
    assert np.allclose(ctx.element_constant("Le"), s.Le, rtol=0, atol=1e-13 * np.abs(s.Le).max())
    for d, k in enumerate(("Gx", "Gy", "Gz")):
        assert np.allclose(ctx.element_constant(k), s.Ge[d], rtol=0, atol=1e-13 * np.abs(s.Ge[d]).max())
    assert np.allclose(_np(ctx.lumped_mass()), s.M, rtol=1e-13)


@pytest.mark.parametrize("beta", [0.5, 1.0])
def test_hex27_convection_matrices(beta):
    print(f"\n[case] hex27 Ce beta={beta}", flush=True)
    _, timestep, s, _, _ = _channel((3, 2, 2), (1.5, 1.0, 0.5))
    U = RNG.standard_normal(3 * s.nn)
    Ce = _np(timestep.context.conv_elem_calc(U, beta))
    ref = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 3, beta)
    assert Ce.shape == ref.shape == (s.nconn.shape[0], 27, 27)
    assert np.abs(Ce - ref).max() <= 1e-13 * np.abs(ref).max()


def test_hex27_face_integral_and_mass_matvec():
    print("\n[case] hex27 face integral, mass matvec", flush=True)
    compute, timestep, s, surf, U_d = _channel((3, 2, 2), (1.5, 1.0, 1.0))
    ctx = timestep.context
    F = _np(ctx.boundary_integral(U_d))
    ref = el.u_dot_n_boundary_integral_quad9_faces(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords)
    assert np.abs(F - ref).max() <= 1e-14 * np.abs(ref).max()
    x = RNG.standard_normal(3 * s.nn)
    assert _rel(_np(ctx.mass_matvec(x)), emult.consistent_mass_mult(x, s.Me, s.nconn, s.nn, 3)) < 1e-14


def test_hex27_pressure_cg_matches_oracle():
    """PCG (approx_inv) on the pinned hex27 Laplacian: the oracle's iteration count and solution."""
    from oracle.solvers import cg
    print("\n[case] hex27 PCG approx_inv", flush=True)
    _, timestep, s, _, _ = _channel((4, 2, 2), precond="approx_inv")
    ctx = timestep.context
    b = RNG.standard_normal(s.nn)
    b[s.gamma_p] = 0
    A = emult.assemble(s.nconn, s.nn, s.Le).tocsr()

    def op(x):
        y = A @ x
        y[s.gamma_p] = 0
        return y
    xo, ko = cg(op, b, tol=1e-10, M=s.M_op)
    x, k, conv = ctx.pressure_cg(b, tol=1e-10)
    print(f"  iters {k} oracle {ko}", flush=True)
    assert conv and k == ko
    assert _rel(_np(x), xo) < 1e-10


@pytest.mark.parametrize("precond", ["jacobi", "approx_inv", None, "multigrid"])
def test_hex27_channel_steps_match_oracle(precond):
    """4 channel steps from rest, FP64: equal BiCGSTAB and PCG iteration counts every step, U and
    P to 1e-9 relative L2.  The bar is looser than the Q1 paths' 1e-10 because this system
    amplifies rounding: in the oracle alone, a 1e-15 relative perturbation of U after the first
    step changes U by 2.3e-10 two steps later with approx_inv (5e-12 with jacobi).  Without a
    preconditioner the first pressure solve stops at its tolerance on a borderline iteration: the
    oracle takes 54 iterations, and 55 after any 1e-15 relative perturbation of its inputs, so the
    pressure count is compared within 1 there."""
    n, dt, rho, mu = (4, 2, 2), 2e-3, 1.0, 1e-3
    print(f"\n[case] hex27 channel {n} precond={precond}", flush=True)
    compute, timestep, s, surf, U_d = _channel(n, precond=precond, final=[2, 1, 1] if precond == "multigrid"
                                               else None)
    U = np.zeros(3 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for k in range(4):
        F_p, F = compute(U_d, None, 0, 0)
        Ug, Pg, ui, pi = timestep(Ug, Pg, dt, rho, mu, U_d, F_p, F)
        oF_p, oF = ns.compute(s, U_d, None, 0, 0)
        U, P, oui, opi = ns.four_step_timestep(s, U, P, dt, rho, mu, U_d, oF_p, oF)
        eu, ep = _rel(_np(Ug), U), _rel(_np(Pg), P)
        print(f"  step {k}: iters {ui},{pi} oracle {oui},{opi}  rel U {eu:.2e} P {ep:.2e}", flush=True)
        assert ui == oui and abs(pi - opi) <= (0 if precond else 1)
        assert eu < 1e-9 and ep < 1e-9


def test_hex27_mixed_form_error():
    """FP32-stored Ce (mixed form): per-field error against the FP64 oracle after 4 steps stays at
    the FP32 rounding level of the convection matrices (tolerance 1e-6)."""
    n, dt, rho, mu = (4, 2, 2), 2e-3, 1.0, 1e-3
    print(f"\n[case] hex27 channel {n} mixed", flush=True)
    compute, timestep, s, surf, U_d = _channel(n, precision="mixed")
    assert timestep.context.ce_format == "f32"
    U = np.zeros(3 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for _ in range(4):
        F_p, F = compute(U_d, None, 0, 0)
        Ug, Pg, _, _ = timestep(Ug, Pg, dt, rho, mu, U_d, F_p, F)
        oF_p, oF = ns.compute(s, U_d, None, 0, 0)
        U, P, _, _ = ns.four_step_timestep(s, U, P, dt, rho, mu, U_d, oF_p, oF)
    Ug = _np(Ug)
    errs = [_rel(Ug[d * s.nn:(d + 1) * s.nn], U[d * s.nn:(d + 1) * s.nn]) for d in range(3)]
    ep = _rel(_np(Pg), P)
    print(f"  rel u {errs[0]:.2e} v {errs[1]:.2e} w {errs[2]:.2e} p {ep:.2e}", flush=True)
    assert max(errs) < 1e-6 and ep < 1e-6


def test_hex27_mixed_multigrid_steps():
    """Mixed form with the hex27 box-grid V-cycle (FP32 Ce, dense FP32 inverse of the pinned
    coarse hex27 operator, flexible PCG): 4 channel steps against the FP64 oracle, per field within
    1e-6, momentum iteration counts equal and pressure counts within one."""
    n, dt, rho, mu = (4, 2, 2), 2e-3, 1.0, 1e-3
    print(f"\n[case] hex27 channel {n} multigrid mixed", flush=True)
    compute, timestep, s, surf, U_d = _channel(n, precond="multigrid", final=[2, 1, 1], precision="mixed")
    assert timestep.context.ce_format == "f32"
    U = np.zeros(3 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for k in range(4):
        F_p, F = compute(U_d, None, 0, 0)
        Ug, Pg, ui, pi = timestep(Ug, Pg, dt, rho, mu, U_d, F_p, F)
        oF_p, oF = ns.compute(s, U_d, None, 0, 0)
        U, P, oui, opi = ns.four_step_timestep(s, U, P, dt, rho, mu, U_d, oF_p, oF)
        print(f"  step {k}: iters {ui},{pi} oracle {oui},{opi}", flush=True)
        assert ui == oui and abs(pi - opi) <= 1
    Ug = _np(Ug)
    errs = [_rel(Ug[d * s.nn:(d + 1) * s.nn], U[d * s.nn:(d + 1) * s.nn]) for d in range(3)]
    ep = _rel(_np(Pg), P)
    print(f"  rel u {errs[0]:.2e} v {errs[1]:.2e} w {errs[2]:.2e} p {ep:.2e}", flush=True)
    assert max(errs) < 1e-6 and ep < 1e-6


def test_hex27_uniform_flow_is_kept():
    """U = (1, 0, 0) imposed on every face of a box stays uniform with P = 0."""
    import jax_fem_cfd.setup.simulation as sim
    print("\n[case] hex27 uniform flow", flush=True)
    n, L = [3, 2, 2], [1.5, 1.0, 1.0]
    cfg = sim.Config(mesh="simple", solver="standard", mesh_config=sim.SimpleMeshConfig(n, L, 2, False),
                     solver_config=sim.StandardSolverConfig([0, 1, 2, 3, 4, 5], [], [], False, 1e-12, True,
                                                            "jacobi"))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    g = np.asarray(gamma_d)[np.asarray(gamma_d) < nn]
    U_d = sim.set_face_velocity(U_d, g, gamma_d, 1.0, nn, 1)
    U0 = np.zeros(3 * nn)
    U0[:nn] = 1.0
    F_p, F = compute(U_d, None, 0, 0)
    U, P, ui, pi = timestep(U0, np.zeros(nn), 1e-2, 1.0, 1e-2, U_d, F_p, F)
    assert np.abs(_np(U) - U0).max() < 1e-11 and np.abs(_np(P)).max() < 1e-11


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_vcycle_h27_matches_oracle(set_zeroP):
    """One V-cycle on hex27 levels (8x4x4 -> 2x1x1 elements, node grids 17x9x9 -> 5x3x3): pinned
    (outlet face, coarse CG to 1e-12) and pure Neumann, against the oracle's V-cycle."""
    from jax_fem_cfd._context import Context
    n, L, nf = [8, 4, 4], [2.0, 1.0, 1.0], [2, 1, 1]
    outlets = [1] if set_zeroP else []
    print(f"\n[case] hex27 V-cycle {n} set_zeroP={set_zeroP}", flush=True)
    s, *_ = ns.build(n, L, False, [0, 2, 3, 4, 5], outlets, 1e-12, set_zeroP, "multigrid", nf, shape_func_ord=2)
    c = Context(n[0], n[1], L[0], L[1], False, tol=1e-12, set_zeroP=set_zeroP, precond="multigrid",
                mg_levels=len(s.levels), gamma_p=s.gamma_p if set_zeroP else None,
                mg_gamma_p=[Lv["gamma_p"] for Lv in s.levels] if set_zeroP else None, nez=n[2], lz=L[2], order=2)
    r = RNG.standard_normal(s.nn)
    if set_zeroP:
        r[s.gamma_p] = 0
    else:
        r -= r.mean()
    ref = s.M_op(r)
    got = _np(c.vcycle(r))
    print(f"  rel {_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) < 1e-11
