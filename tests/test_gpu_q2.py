"""GPU parity of the quad9 (Q2) path against the oracle: element constants, convection matrices,
the lin3 boundary integral, the mass matvec and whole four-step timesteps (lid-driven cavity with
each pressure preconditioner, periodic Taylor-Green, the mixed form)."""
import numpy as np
import pytest

from oracle import elements as el
from oracle import emult
from oracle import mesh as ms
from oracle import ns

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)


def _np(x):
    return x.detach().cpu().numpy()


def _rel(a, b):
    return ns.rel_l2(a, b)


def _setup(n=(8, 8), L=(1.0, 1.0), periodic=False, faces=(0, 1, 2, 3), outlets=(), precond="jacobi", tol=1e-10,
           set_zeroP=True, precision="fp64", final=None):
    import jax_fem_cfd.setup.simulation as sim
    cfg = sim.Config(mesh="simple", solver="standard",
                     mesh_config=sim.SimpleMeshConfig(list(n), list(L), 2, periodic),
                     solver_config=sim.StandardSolverConfig(list(faces), [], list(outlets), False, tol, set_zeroP,
                                                            precond, final, precision))
    out = sim.setup(cfg)
    s, surf, _, U_d = ns.build(list(n), list(L), periodic, list(faces), list(outlets), tol, set_zeroP, precond, final,
                               shape_func_ord=2)
    compute, timestep, coords, osurf, h, nn, gamma_d, _ = out
    assert nn == s.nn and np.allclose(coords, s.coords)
    if gamma_d is not None:
        assert np.array_equal(np.asarray(gamma_d), s.gamma_d)
    return compute, timestep, s, surf, U_d


def test_quad9_constants_match_oracle():
    print("\n[case] quad9 constants", flush=True)
    _, timestep, s, _, _ = _setup((4, 3), (2.0, 1.5))
    ctx = timestep.context
    assert ctx.order == 2 and ctx.nn == s.nn
    assert np.allclose(ctx.element_constant("Me"), s.Me, rtol=1e-13, atol=1e-17)
    assert np.allclose(ctx.element_constant("Le"), s.Le, rtol=1e-12, atol=1e-14)
    assert np.allclose(ctx.element_constant("Gx"), s.Ge[0], rtol=1e-12, atol=1e-16)
    assert np.allclose(ctx.element_constant("Gy"), s.Ge[1], rtol=1e-12, atol=1e-16)
    assert np.allclose(_np(ctx.lumped_mass()), s.M, rtol=1e-13)


@pytest.mark.parametrize("periodic", [False, True])
def test_quad9_convection_matrices(periodic):
    print(f"\n[case] quad9 Ce periodic={periodic}", flush=True)
    faces = () if periodic else (0, 1, 2, 3)
    _, timestep, s, _, _ = _setup((5, 3), (2.0, 1.5), periodic, faces, set_zeroP=not periodic)
    U = RNG.standard_normal(2 * s.nn)
    for beta in (0.5, 1.0):
        Ce = _np(timestep.context.conv_elem_calc(U, beta))
        ref = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 2, beta)
        assert np.abs(Ce - ref).max() <= 1e-13 * np.abs(ref).max()


def test_quad9_boundary_integral_and_mass():
    print("\n[case] quad9 F_p and M_c", flush=True)
    compute, timestep, s, surf, U_d = _setup((6, 4), (1.5, 1.0), faces=(0, 2, 3), outlets=(1,))
    U_d = ms.set_face_velocity(U_d, surf[3], s.gamma_d, RNG.standard_normal(len(surf[3])), s.nn, 1)
    U_d = ms.set_face_velocity(U_d, surf[0], s.gamma_d, RNG.standard_normal(len(surf[0])), s.nn, 2)
    F_p = _np(timestep.context.boundary_integral(U_d))
    ref = el.u_dot_n_boundary_integral_1d(s.bconn, s.nvec, s.gamma_d, U_d, s.nn, s.coords, el.lin3_shape_functions)
    assert np.abs(F_p - ref).max() <= 1e-14 * np.abs(ref).max()
    x = RNG.standard_normal(2 * s.nn)
    assert _rel(_np(timestep.context.mass_matvec(x)), emult.consistent_mass_mult(x, s.Me, s.nconn, s.nn, 2)) < 1e-14


@pytest.mark.parametrize("precond", ["jacobi", "approx_inv", None])
def test_quad9_cavity_steps_match_oracle(precond):
    n, dt, rho, mu = (8, 8), 1e-3, 1.0, 1e-2
    print(f"\n[case] quad9 cavity {n} precond={precond}", flush=True)
    compute, timestep, s, surf, U_d = _setup(n, precond=precond)
    U_d = ns.cavity_lid(s, surf, U_d)
    U = np.zeros(2 * s.nn)
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


def test_quad9_taylor_green_periodic_matches_oracle():
    n, dt, rho, mu = (8, 8), 1e-2, 1.0, 0.01
    print(f"\n[case] quad9 Taylor-Green {n}", flush=True)
    compute, timestep, s, surf, _ = _setup(n, (2 * np.pi, 2 * np.pi), True, (), precond="jacobi", set_zeroP=False)
    u, v, p = ns.taylor_green(s.coords, 0.0, rho, mu)
    U, P = np.concatenate([u, v]), p
    Ug, Pg = U.copy(), P.copy()
    for _ in range(3):
        Ug, Pg, ui, pi = timestep(Ug, Pg, dt, rho, mu, None, 0, 0)
        U, P, oui, opi = ns.four_step_timestep(s, U, P, dt, rho, mu, None, 0, 0)
        assert (ui, pi) == (oui, opi)
    assert _rel(_np(Ug), U) < 1e-10 and _rel(_np(Pg), P) < 1e-10


def test_quad9_mixed_form_error():
    """FP32-stored Ce: per-field error against the FP64 oracle after 4 cavity steps < 1e-6."""
    n, dt, rho, mu = (8, 8), 1e-3, 1.0, 1e-2
    print(f"\n[case] quad9 cavity {n} mixed", flush=True)
    compute, timestep, s, surf, U_d = _setup(n, precision="mixed")
    U_d = ns.cavity_lid(s, surf, U_d)
    U = np.zeros(2 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for _ in range(4):
        F_p, F = compute(U_d, None, 0, 0)
        Ug, Pg, _, _ = timestep(Ug, Pg, dt, rho, mu, U_d, F_p, F)
        oF_p, oF = ns.compute(s, U_d, None, 0, 0)
        U, P, _, _ = ns.four_step_timestep(s, U, P, dt, rho, mu, U_d, oF_p, oF)
    Ug = _np(Ug)
    eu, ev, ep = _rel(Ug[:s.nn], U[:s.nn]), _rel(Ug[s.nn:], U[s.nn:]), _rel(_np(Pg), P)
    print(f"  rel u {eu:.2e} v {ev:.2e} p {ep:.2e}", flush=True)
    assert eu < 1e-6 and ev < 1e-6 and ep < 1e-6


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_quad9_vcycle_matches_oracle(set_zeroP):
    """One quad9 V-cycle 16x16 -> 2x2 (4 levels), pinned and pure Neumann, against the oracle."""
    from oracle import multigrid as mg
    print(f"\n[case] quad9 V-cycle set_zeroP={set_zeroP}", flush=True)
    compute, timestep, s, surf, U_d = _setup((16, 16), precond="multigrid", set_zeroP=set_zeroP, final=[2, 2])
    r = RNG.standard_normal(s.nn)
    if not set_zeroP:
        r -= r.mean()
    z = _np(timestep.context.vcycle(r))
    ref = mg.v_cycle(r, s.levels, s.tol, False, set_zeroP)
    print(f"  rel {_rel(z, ref):.3e}", flush=True)
    # the pinned cycle ends in a coarse CG to tol whose rounding follows the summation order
    assert _rel(z, ref) < 1e-11


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_quad9_cavity_multigrid_steps_match_oracle(set_zeroP):
    n, dt, rho, mu = (16, 16), 1e-3, 1.0, 1e-2
    print(f"\n[case] quad9 cavity {n} multigrid set_zeroP={set_zeroP}", flush=True)
    compute, timestep, s, surf, U_d = _setup(n, precond="multigrid", set_zeroP=set_zeroP, final=[2, 2])
    U_d = ns.cavity_lid(s, surf, U_d)
    U = np.zeros(2 * s.nn)
    U[s.gamma_d] = U_d
    P = np.zeros(s.nn)
    Ug, Pg = U.copy(), P.copy()
    for k in range(3):
        F_p, F = compute(U_d, None, 0, 0)
        Ug, Pg, ui, pi = timestep(Ug, Pg, dt, rho, mu, U_d, F_p, F)
        oF_p, oF = ns.compute(s, U_d, None, 0, 0)
        U, P, oui, opi = ns.four_step_timestep(s, U, P, dt, rho, mu, U_d, oF_p, oF)
        eu, ep = _rel(_np(Ug), U), _rel(_np(Pg), P)
        print(f"  step {k}: iters {ui},{pi} oracle {oui},{opi}  rel U {eu:.2e} P {ep:.2e}", flush=True)
        assert (ui, pi) == (oui, opi)
        assert eu < 1e-10 and ep < 1e-10


def test_quad9_taylor_green_multigrid_matches_oracle():
    n, dt, rho, mu = (16, 16), 1e-2, 1.0, 0.01
    print(f"\n[case] quad9 Taylor-Green {n} multigrid", flush=True)
    compute, timestep, s, surf, _ = _setup(n, (2 * np.pi, 2 * np.pi), True, (), precond="multigrid",
                                           set_zeroP=False, final=[2, 2])
    u, v, p = ns.taylor_green(s.coords, 0.0, rho, mu)
    U, P = np.concatenate([u, v]), p
    Ug, Pg = U.copy(), P.copy()
    for _ in range(3):
        Ug, Pg, ui, pi = timestep(Ug, Pg, dt, rho, mu, None, 0, 0)
        U, P, oui, opi = ns.four_step_timestep(s, U, P, dt, rho, mu, None, 0, 0)
        assert (ui, pi) == (oui, opi)
    assert _rel(_np(Ug), U) < 1e-10 and _rel(_np(Pg), P) < 1e-10


# ----------------------------------------------------------------------------------------- passive scalar
def _passive_setup(n, precond, supg, phi_faces=(), final=None):
    import jax_fem_cfd.setup.simulation as sim
    cfg = sim.Config(mesh="simple", solver="passive_transport",
                     mesh_config=sim.SimpleMeshConfig(list(n), [1.0, 1.0], 2, False),
                     solver_config=sim.PassiveTransportConfig(
                         sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, 1e-10, True, precond, final),
                         sim.ADEConfig(list(phi_faces), supg)))
    compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi = sim.setup(cfg)
    s, osurf, _, oUd = ns.build(list(n), [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-10, True, precond, final,
                                shape_func_ord=2, passive=True)
    assert nn == s.nn and np.allclose(coords, s.coords)
    return compute, timestep, coords, gamma_d, U_d, gamma_d_phi, s, oUd


@pytest.mark.parametrize("precond,supg,phi_faces", [("jacobi", False, ()), ("jacobi", True, ()),
                                                    ("multigrid", True, (3,)), ("approx_inv", False, (3,))])
def test_quad9_passive_transport_matches_oracle(precond, supg, phi_faces):
    """Passive scalar on quad9 (passive_transport.py:64-97, 114-116): swirl in a cavity, scalar
    rectangle, optional phi = 1 on the left wall; U, P, phi against the oracle with equal
    iteration counts.  approx_inv uses the passive solver's closure (:154-162)."""
    n, dt, rho, mu, D = (8, 8), 2e-3, 1.0, 1e-2, 1e-4
    final = [2, 2] if precond == "multigrid" else None
    print(f"\n[case] quad9 passive {n} precond={precond} supg={supg} phi_faces={phi_faces}", flush=True)
    compute, timestep, coords, gamma_d, U_d, gamma_d_phi, s, oUd = _passive_setup(n, precond, supg, phi_faces,
                                                                                  final)
    x, y = coords[:, 0], coords[:, 1]
    U0 = np.concatenate([np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y)])
    U0[gamma_d] = U_d
    phi0 = ((x >= 0.3) & (x <= 0.7) & (y >= 0.2) & (y <= 0.6)).astype(float)
    phi_d = None if gamma_d_phi is None else np.ones(len(gamma_d_phi))
    U, P, phi = U0, np.zeros(s.nn), phi0
    Uo, Po, phio = U0.copy(), np.zeros(s.nn), phi0.copy()
    for k in range(3):
        F_p, F = compute(U_d, None, 0, 0)
        U, P, phi, ui, pi, fi = timestep(U, P, phi, dt, rho, mu, D, U_d, phi_d, F_p, F, None)
        Uo, Po, phio, uio, pio, fio = ns.passive_timestep(s, Uo, Po, phio, dt, rho, mu, D, oUd, phi_d, 0, 0,
                                                          gamma_d_phi, supg)
        ephi = _rel(_np(phi), phio)
        print(f"  step {k}: iters {ui},{pi},{fi} oracle {uio},{pio},{fio}  rel phi {ephi:.2e}", flush=True)
        assert (ui, pi, fi) == (uio, pio, fio)
    assert _rel(_np(U), Uo) < 1e-10 and _rel(_np(P), Po) < 1e-10 and _rel(_np(phi), phio) < 1e-10
    if gamma_d_phi is not None:
        assert np.array_equal(_np(phi)[gamma_d_phi], phi_d)


# ----------------------------------------------------------------------------------------- reverse mode
def _vjp_state(s, U_d, seed=5):
    r = np.random.default_rng(seed)
    x, y = s.coords[:, 0], s.coords[:, 1]
    U = np.concatenate([np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y)])
    U = U + 0.05 * r.standard_normal(U.size)
    U[s.gamma_d] = U_d
    return U, 0.1 * r.standard_normal(s.nn)


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_quad9_ns_step_vjp_matches_oracle(set_zeroP):
    """Reverse-mode product of one quad9 NS step (lid-driven cavity, multigrid PCG; pinned and
    pure-Neumann pressure) against the oracle's VJP, itself pinned by finite differences
    (tests/test_adjoint_oracle.py): cotangents to 1e-9, forward and pressure-adjoint iteration
    counts equal, the momentum-adjoint BiCGSTAB count within 2 (its start is a random cotangent,
    where rounding decides the last iteration)."""
    from oracle import adjoint as oadj
    n, dt, rho, mu = (8, 8), 2e-3, 1.0, 1e-2
    print(f"\n[case] quad9 NS step VJP {n} set_zeroP={set_zeroP}", flush=True)
    compute, timestep, s, surf, U_d = _setup(n, precond="multigrid", set_zeroP=set_zeroP, final=[2, 2])
    U_d = ns.cavity_lid(s, surf, U_d)
    U_n, P_n = _vjp_state(s, U_d)
    gU, gP = RNG.standard_normal(2 * s.nn), RNG.standard_normal(s.nn)
    F_p, F = compute(U_d, None, 0, 0)
    ctx = timestep.context
    (gUn, gPn), (U, P), its = ctx.four_step_timestep_vjp(U_n, P_n, dt, rho, mu, gU, gP, U_d, F_p, F)
    oF_p, _ = ns.compute(s, U_d, None, 0, 0)
    ogUn, ogPn, fw, (m_it, p_it) = oadj.four_step_vjp(s, U_n, P_n, dt, rho, mu, U_d, oF_p, 0, gU, gP)
    print(f"  iters {its} oracle {(fw['U_it'], fw['P_it'], p_it, m_it)}  rel gU_n {_rel(_np(gUn), ogUn):.3e}"
          f"  gP_n {_rel(_np(gPn), ogPn):.3e}", flush=True)
    assert _rel(_np(U), fw["U"]) < 1e-10 and _rel(_np(P), fw["P"]) < 1e-10
    assert _rel(_np(gUn), ogUn) < 1e-9 and _rel(_np(gPn), ogPn) < 1e-9
    assert its[:3] == (fw["U_it"], fw["P_it"], p_it) and abs(its[3] - m_it) <= 2


def test_quad9_rollout_gradient_matches_finite_differences_on_gpu():
    """Gradient of L = a . U_N + b . P_N after 6 quad9 steps (checkpoint 2) through the
    differentiable rollout (torch autograd over the adjoint sweep) against central differences
    of the GPU forward rollout, solves at 1e-12."""
    import torch
    from jax_fem_cfd.solvers import adjoint as hadj
    n, nsteps, dt, rho, mu = (8, 8), 6, 2e-3, 1.0, 1e-2
    print(f"\n[case] quad9 rollout gradient vs FD {n} steps={nsteps}", flush=True)
    compute, timestep, s, surf, U_d = _setup(n, precond="multigrid", tol=1e-12, final=[2, 2])
    U_d = ns.cavity_lid(s, surf, U_d)
    U0, P0 = _vjp_state(s, U_d)
    ctx = timestep.context
    F_p, F = compute(U_d, None, 0, 0)
    dev = torch.device("cuda", 0)
    a = torch.tensor(RNG.standard_normal(2 * s.nn), dtype=torch.float64, device=dev)
    b = torch.tensor(RNG.standard_normal(s.nn), dtype=torch.float64, device=dev)
    P0t = torch.tensor(P0, dtype=torch.float64, device=dev)

    def loss_of(U):
        Uf, Pf, _ = hadj.rollout(ctx, nsteps, U, P0t, None, dt, rho, mu, U_d=U_d, F_p=F_p, F=F, checkpoint=2)
        return torch.dot(a, Uf) + torch.dot(b, Pf)

    U0t = torch.tensor(U0, dtype=torch.float64, device=dev, requires_grad=True)
    loss_of(U0t).backward()
    g = U0t.grad.detach()
    with torch.no_grad():
        for seed in range(2):
            d = torch.tensor(np.random.default_rng(seed).standard_normal(2 * s.nn), dtype=torch.float64, device=dev)
            d[torch.as_tensor(s.gamma_d, device=dev)] = 0.0
            eps = 1e-4
            fd = (loss_of(U0t.detach() + eps * d) - loss_of(U0t.detach() - eps * d)).item() / (2 * eps)
            an = torch.dot(g, d).item()
            print(f"  FD {fd:.12e}  adjoint {an:.12e}", flush=True)
            assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


@pytest.mark.parametrize("phi_faces,supg", [((), False), ((3,), False), ((), True), ((3,), True)])
def test_quad9_passive_step_vjp_matches_oracle(phi_faces, supg):
    """Reverse-mode product of one passive quad9 step, plain and SUPG, against the oracle's VJP
    (pinned by finite differences in tests/test_adjoint_oracle.py): cotangents to 1e-9; the
    symmetric pressure-adjoint CG count equal, the transposed BiCGSTAB counts within 2."""
    from oracle import adjoint as oadj
    n, dt, rho, mu, D = (8, 8), 2e-3, 1.0, 1e-2, 1e-3
    print(f"\n[case] quad9 passive step VJP {n} phi_faces={phi_faces} supg={supg}", flush=True)
    compute, timestep, coords, gamma_d, U_d, gamma_d_phi, s, oUd = _passive_setup(n, "multigrid", supg, phi_faces,
                                                                                  [2, 2])
    U_d = ns.cavity_lid(s, [np.asarray(x) for x in ms.create_mesh_2d_quad9(8, 8, 1.0, 1.0)[2]], U_d)
    U_n, P_n = _vjp_state(s, U_d)
    x, y = s.coords[:, 0], s.coords[:, 1]
    phi_n = ((x > 0.2) & (x < 0.6) & (y > 0.3) & (y < 0.7)).astype(float)
    phi_d = None if gamma_d_phi is None else np.linspace(0.0, 1.0, len(gamma_d_phi))
    gU, gP, gphi = RNG.standard_normal(2 * s.nn), RNG.standard_normal(s.nn), RNG.standard_normal(s.nn)
    ctx = timestep.context
    F_p, F = compute(U_d, None, 0, 0)
    (gUn, gPn, gfn), (U, P, phi), its = ctx.passive_timestep_vjp(U_n, P_n, phi_n, dt, rho, mu, D, gU, gP, gphi,
                                                                 U_d, phi_d, F_p, F, supg)
    oF_p, _ = ns.compute(s, U_d, None, 0, 0)
    ogUn, ogPn, ogfn, (oU, oP, ophi), oits = oadj.passive_vjp(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, phi_d, oF_p,
                                                              0, gamma_d_phi, gU, gP, gphi, supg=supg)
    print(f"  iters {its[3:]} oracle {oits}  rel " + " ".join(
        f"{_rel(_np(a), b):.3e}" for a, b in ((gUn, ogUn), (gPn, ogPn), (gfn, ogfn))), flush=True)
    assert _rel(_np(phi), ophi) < 1e-10 and _rel(_np(U), oU) < 1e-10
    for a, b in ((gUn, ogUn), (gPn, ogPn), (gfn, ogfn)):
        assert _rel(_np(a), b) < 1e-9, _rel(_np(a), b)
    assert its[4] == oits[1] and abs(its[3] - oits[0]) <= 2 and abs(its[5] - oits[2]) <= 2
