"""GPU parity of the 3D (hex8) path against the oracle: element constants, convection
matrices, the face boundary integral, the mass matvec, the box-grid V-cycle and whole
four-step timesteps of a channel (parabolic inflow, outlet, no-slip walls) with each 3D
pressure preconditioner, and of a closed lid-driven box (pure-Neumann pressure)."""
import numpy as np
import pytest

from oracle import elements as el
from oracle import emult
from oracle import mesh as ms
from oracle import ns

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(77)


def _np(x):
    return x.detach().cpu().numpy()


def _rel(a, b):
    return ns.rel_l2(a, b)


def _channel(n=(8, 4, 4), L=(2.0, 1.0, 1.0), precond="jacobi", tol=1e-10, precision="fp64", final=None):
    import jax_fem_cfd.setup.simulation as sim
    cfg = sim.Config(mesh="simple", solver="standard",
                     mesh_config=sim.SimpleMeshConfig(list(n), list(L), 1, False),
                     solver_config=sim.StandardSolverConfig([0, 2, 3, 4, 5], [], [1], False, tol, True, precond,
                                                            final, precision))
    out = sim.setup(cfg)
    s, surf, _, _ = ns.build(list(n), list(L), False, [0, 2, 3, 4, 5], [1], tol, True, precond, final)
    compute, timestep, coords, osurf, h, nn, gamma_d, U_d = out
    assert np.array_equal(np.asarray(gamma_d), s.gamma_d)
    U_d = ns.channel_inflow(s, surf, np.zeros(len(s.gamma_d)), L[1], L[2])
    return compute, timestep, s, surf, U_d


def test_hex8_constants_match_oracle():
    print("\n[case] hex8 element constants", flush=True)
    _, timestep, s, _, _ = _channel((4, 3, 2), (2.0, 1.5, 1.0))
    ctx = timestep.context
    assert ctx.ndim == 3 and ctx.nn == s.nn
    assert np.allclose(ctx.element_constant("Me"), s.Me, rtol=1e-14, atol=0)
    assert np.allclose(ctx.element_constant("Le"), s.Le, rtol=1e-13, atol=1e-16)
    for d, k in enumerate(("Gx", "Gy", "Gz")):
        assert np.allclose(ctx.element_constant(k), s.Ge[d], rtol=1e-13, atol=1e-18)
    assert np.allclose(_np(ctx.lumped_mass()), s.M, rtol=1e-14)


@pytest.mark.parametrize("beta", [0.5, 1.0])
def test_hex8_convection_matrices(beta):
    print(f"\n[case] hex8 Ce beta={beta}", flush=True)
    _, timestep, s, _, _ = _channel((5, 3, 2), (2.0, 1.5, 1.0))
    U = RNG.standard_normal(3 * s.nn)
    Ce = _np(timestep.context.conv_elem_calc(U, beta))
    ref = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 3, beta)
    assert Ce.shape == ref.shape
    assert np.abs(Ce - ref).max() <= 1e-13 * np.abs(ref).max()


def test_face_integral_and_mass_matvec():
    print("\n[case] hex8 F_p and M_c", flush=True)
    compute, timestep, s, surf, U_d = _channel((6, 4, 3), (2.0, 1.0, 0.75))
    ctx = timestep.context
    F_p = _np(ctx.boundary_integral(U_d))
    ref = el.u_dot_n_boundary_integral_2d(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords)
    assert np.abs(F_p - ref).max() <= 1e-14 * np.abs(ref).max() + 1e-18
    assert np.isclose(F_p.sum(), ref.sum())
    x = RNG.standard_normal(3 * s.nn)
    y = _np(ctx.mass_matvec(x))
    assert _rel(y, emult.consistent_mass_mult(x, s.Me, s.nconn, s.nn, 3)) < 1e-14
    # compute() with a source term: F = M_c (S(t_n) + S(t)) / 2 (standard.py:184-186)
    S = RNG.standard_normal(3 * s.nn)
    _, F = compute(U_d, lambda t: S * (1 + t), 0.0, 1.0)
    _, Fo = ns.compute(s, U_d, lambda t: S * (1 + t), 0.0, 1.0)
    assert _rel(_np(F), Fo) < 1e-14


@pytest.mark.parametrize("precond", ["jacobi", "approx_inv", None])
def test_channel_steps_match_oracle(precond):
    """5 channel steps from rest, FP64: velocity and pressure to 1e-10 relative L2, equal
    BiCGSTAB and PCG iteration counts every step."""
    n, dt, rho, mu = (8, 4, 4), 2e-3, 1.0, 1e-3
    print(f"\n[case] hex8 channel {n} precond={precond}", flush=True)
    compute, timestep, s, surf, U_d = _channel(n, precond=precond)
    U = np.zeros(3 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for k in range(5):
        F_p, F = compute(U_d, None, 0, 0)
        Ug, Pg, ui, pi = timestep(Ug, Pg, dt, rho, mu, U_d, F_p, F)
        oF_p, oF = ns.compute(s, U_d, None, 0, 0)
        U, P, oui, opi = ns.four_step_timestep(s, U, P, dt, rho, mu, U_d, oF_p, oF)
        eu, ep = _rel(_np(Ug), U), _rel(_np(Pg), P)
        print(f"  step {k}: iters {ui},{pi} oracle {oui},{opi}  rel U {eu:.2e} P {ep:.2e}", flush=True)
        assert (ui, pi) == (oui, opi)
        assert eu < 1e-10 and ep < 1e-10


def test_channel_mixed_form_error():
    """FP32-stored Ce (mixed form): per-field error (u, v, w, p) against the FP64 oracle after 5
    steps stays at the FP32 rounding level of the convection matrices (tolerance 1e-6 each)."""
    n, dt, rho, mu = (8, 4, 4), 2e-3, 1.0, 1e-3
    print(f"\n[case] hex8 channel {n} mixed", flush=True)
    compute, timestep, s, surf, U_d = _channel(n, precision="mixed")
    assert timestep.context.ce_format == "f32"
    U = np.zeros(3 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for _ in range(5):
        F_p, F = compute(U_d, None, 0, 0)
        Ug, Pg, _, _ = timestep(Ug, Pg, dt, rho, mu, U_d, F_p, F)
        oF_p, oF = ns.compute(s, U_d, None, 0, 0)
        U, P, _, _ = ns.four_step_timestep(s, U, P, dt, rho, mu, U_d, oF_p, oF)
    Ug = _np(Ug)
    errs = [_rel(Ug[d * s.nn:(d + 1) * s.nn], U[d * s.nn:(d + 1) * s.nn]) for d in range(3)]
    ep = _rel(_np(Pg), P)
    print(f"  rel u {errs[0]:.2e} v {errs[1]:.2e} w {errs[2]:.2e} p {ep:.2e}", flush=True)
    assert max(errs) < 1e-6 and ep < 1e-6


def test_uniform_flow_is_kept():
    """U = (1, 0, 0) imposed on every face of a box stays uniform with P = 0."""
    import jax_fem_cfd.setup.simulation as sim
    print("\n[case] hex8 uniform flow", flush=True)
    n, L = [4, 3, 2], [2.0, 1.5, 1.0]
    cfg = sim.Config(mesh="simple", solver="standard", mesh_config=sim.SimpleMeshConfig(n, L, 1, False),
                     solver_config=sim.StandardSolverConfig([0, 1, 2, 3, 4, 5], [], [], False, 1e-12, True,
                                                            "jacobi"))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    g = np.asarray(gamma_d)[np.asarray(gamma_d) < nn]
    U_d = sim.set_face_velocity(U_d, g, gamma_d, 1.0, nn, 1)
    U0 = np.zeros(3 * nn)
    U0[:nn] = 1.0
    F_p, F = compute(U_d, None, 0, 0)
    U, P, ui, pi = timestep(U0, np.zeros(nn), 1e-2, 1.0, 1e-2, U_d, F_p, F)
    assert np.abs(_np(U) - U0).max() < 1e-12 and np.abs(_np(P)).max() < 1e-12


def test_2d_only_entry_points_refuse_3d_contexts():
    from jax_fem_cfd._native import NativeError
    print("\n[case] hex8 refusals", flush=True)
    _, timestep, s, _, _ = _channel((4, 2, 2), (2.0, 1.0, 1.0))
    ctx = timestep.context
    with pytest.raises(NativeError):
        ctx.vcycle(np.zeros(s.nn))
    with pytest.raises(NativeError):
        ctx.four_step_timestep_vjp(np.zeros(3 * s.nn), np.zeros(s.nn), 1e-3, 1.0, 1e-3, np.zeros(3 * s.nn),
                                   np.zeros(s.nn))


# ----------------------------------------------------------------------------------------- multigrid
@pytest.mark.parametrize("set_zeroP", [True, False])
def test_vcycle3_matches_oracle(set_zeroP):
    """One box-grid V-cycle (4 levels, 16x8x8 -> 2x1x1) on a random residual: pinned (outlet
    face, coarse CG to 1e-12) and pure Neumann (closed box), against the oracle's V-cycle."""
    from jax_fem_cfd._context import Context
    n, L, nf = [16, 8, 8], [2.0, 1.0, 1.0], [2, 1, 1]
    outlets = [1] if set_zeroP else []
    print(f"\n[case] hex8 V-cycle {n} set_zeroP={set_zeroP}", flush=True)
    s, *_ = ns.build(n, L, False, [0, 2, 3, 4, 5], outlets, 1e-12, set_zeroP, "multigrid", nf)
    c = Context(n[0], n[1], L[0], L[1], False, tol=1e-12, set_zeroP=set_zeroP, precond="multigrid",
                mg_levels=len(s.levels), gamma_p=s.gamma_p if set_zeroP else None,
                mg_gamma_p=[Lv["gamma_p"] for Lv in s.levels] if set_zeroP else None, nez=n[2], lz=L[2])
    r = RNG.standard_normal(s.nn)
    if set_zeroP:
        r[s.gamma_p] = 0
    else:
        r -= r.mean()
    ref = s.M_op(r)
    got = _np(c.vcycle(r))
    print(f"  rel {_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) < 1e-11


def test_pressure_cg3_multigrid():
    """PCG with the box-grid V-cycle on the pinned channel Laplacian: converged to 1e-8 in the
    oracle's iteration count, with the oracle's solution."""
    from jax_fem_cfd._context import Context
    from oracle.solvers import cg
    n, L, nf = [16, 8, 8], [2.0, 1.0, 1.0], [4, 2, 2]
    print(f"\n[case] hex8 MG-PCG {n}", flush=True)
    s, *_ = ns.build(n, L, False, [0, 2, 3, 4, 5], [1], 1e-8, True, "multigrid", nf)
    c = Context(n[0], n[1], L[0], L[1], False, tol=1e-8, set_zeroP=True, precond="multigrid",
                mg_levels=len(s.levels), gamma_p=s.gamma_p, mg_gamma_p=[Lv["gamma_p"] for Lv in s.levels],
                nez=n[2], lz=L[2])
    b = RNG.standard_normal(s.nn)
    b[s.gamma_p] = 0
    A = emult.assemble(s.nconn, s.nn, s.Le).tocsr()

    def op(x):
        y = A @ x
        y[s.gamma_p] = 0
        return y
    xo, ko = cg(op, b, tol=1e-8, M=s.M_op)
    x, k, conv = c.pressure_cg(b, tol=1e-8)
    print(f"  iters {k} oracle {ko}", flush=True)
    assert conv and k == ko
    assert _rel(_np(x), xo) < 1e-10


def test_channel_multigrid_steps_match_oracle():
    """5 channel steps from rest with the box-grid V-cycle (16x8x8 -> 2x1x1, outlet pins on
    every level), FP64: U and P to 1e-10, equal iteration counts every step."""
    n, dt, rho, mu = (16, 8, 8), 2e-3, 1.0, 1e-3
    print(f"\n[case] hex8 channel {n} multigrid", flush=True)
    compute, timestep, s, surf, U_d = _channel(n, precond="multigrid", final=[2, 1, 1])
    assert timestep.context.precond == "multigrid"
    U = np.zeros(3 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for k in range(5):
        F_p, F = compute(U_d, None, 0, 0)
        Ug, Pg, ui, pi = timestep(Ug, Pg, dt, rho, mu, U_d, F_p, F)
        oF_p, oF = ns.compute(s, U_d, None, 0, 0)
        U, P, oui, opi = ns.four_step_timestep(s, U, P, dt, rho, mu, U_d, oF_p, oF)
        eu, ep = _rel(_np(Ug), U), _rel(_np(Pg), P)
        print(f"  step {k}: iters {ui},{pi} oracle {oui},{opi}  rel U {eu:.2e} P {ep:.2e}", flush=True)
        assert (ui, pi) == (oui, opi)
        assert eu < 1e-10 and ep < 1e-10


def test_vcycle3_mixed_dense_coarse():
    """Mixed form, pinned: the coarse solve is the exact inverse of the pinned coarse operator
    (stored FP32).  One V-cycle equals the oracle's (coarse CG to 1e-12) to the FP32 rounding of
    that inverse (tolerance 1e-6), and far from a V-cycle whose coarse solve is skipped."""
    from jax_fem_cfd._context import Context
    n, L, nf = [16, 8, 8], [2.0, 1.0, 1.0], [2, 1, 1]
    print(f"\n[case] hex8 V-cycle {n} mixed, dense coarse inverse", flush=True)
    s, *_ = ns.build(n, L, False, [0, 2, 3, 4, 5], [1], 1e-12, True, "multigrid", nf)
    c = Context(n[0], n[1], L[0], L[1], False, tol=1e-12, set_zeroP=True, precond="multigrid",
                mg_levels=len(s.levels), gamma_p=s.gamma_p, mg_gamma_p=[Lv["gamma_p"] for Lv in s.levels],
                nez=n[2], lz=L[2], precision="mixed")
    r = RNG.standard_normal(s.nn)
    r[s.gamma_p] = 0
    ref = s.M_op(r)
    got = _np(c.vcycle(r))
    print(f"  rel {_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) < 1e-6


def test_channel_mixed_multigrid_steps():
    """Mixed form with the box-grid V-cycle (FP32 Ce, dense FP32 coarse inverse, flexible PCG):
    5 channel steps from rest against the FP64 oracle, per field (u, v, w, p) within 1e-6, the
    momentum iteration counts equal and the pressure counts within one."""
    n, dt, rho, mu = (16, 8, 8), 2e-3, 1.0, 1e-3
    print(f"\n[case] hex8 channel {n} multigrid mixed", flush=True)
    compute, timestep, s, surf, U_d = _channel(n, precond="multigrid", final=[2, 1, 1], precision="mixed")
    assert timestep.context.ce_format == "f32"
    U = np.zeros(3 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for k in range(5):
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


def test_lid_box_neumann_multigrid_steps_match_oracle():
    """Closed box, lid u = 1 on the top face, pure-Neumann pressure (no pins) with the box-grid
    V-cycle: 4 steps, U and P (mean-projected) to 1e-10, equal iteration counts."""
    import jax_fem_cfd.setup.simulation as sim
    n, L, dt, rho, mu = [8, 8, 8], [1.0, 1.0, 1.0], 2e-3, 1.0, 1e-2
    print(f"\n[case] hex8 lid box {n} Neumann multigrid", flush=True)
    faces = [0, 1, 2, 3, 4, 5]
    cfg = sim.Config(mesh="simple", solver="standard", mesh_config=sim.SimpleMeshConfig(n, L, 1, False),
                     solver_config=sim.StandardSolverConfig(faces, [], [], False, 1e-10, False, "multigrid", [2, 2, 2]))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    s, osurf, _, oUd = ns.build(n, L, False, faces, [], 1e-10, False, "multigrid", [2, 2, 2])
    top = np.setdiff1d(osurf[2], np.concatenate([osurf[f] for f in (0, 1, 3, 4, 5)]))
    U_d = ms.set_face_velocity(oUd, top, s.gamma_d, 1.0, s.nn, 1)
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
        assert (ui, pi) == (oui, opi)
        assert eu < 1e-10 and ep < 1e-10
