"""GPU parity of the reverse-mode products (fcfd_*_vjp) and of the differentiable
rollout (timestep_loop_scan + torch autograd) against the oracle adjoint
(oracle/adjoint.py, itself pinned by finite differences on the CPU), and a
central finite-difference check of the GPU gradient itself.

Tolerances: VJPs and rollout gradients 1e-9 relative L2 against the oracle (the
forward step agrees to 1e-10; the adjoint solves run the same algorithms at the
same tolerance); the FD check 1e-5 relative on a directional derivative, with the
solves at 1e-12.  Iteration counts: forward solves and the adjoint CG equal; the
transposed BiCGSTAB solves within one iteration.  Those start, as the reference's
transposed solve does, from the forward solve's x0 (a velocity / scalar field, not
an estimate of the adjoint), so their residual falls from |A^T x0| >> |b| to
tol |b| and the last iterations run where the rounding of the two summation
orders differs.
"""
import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import ns

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(424242)


def _np(x):
    return x.detach().cpu().numpy()


def _rel(a, b):
    return ns.rel_l2(a, b)


def _setup(n, solver, set_zeroP=True, tol=1e-10, phi_faces=(), final=(4, 4)):
    import jax_fem_cfd.setup.simulation as sim
    ns_cfg = sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, tol, set_zeroP, "multigrid", list(final))
    if solver == "standard":
        cfg = sim.Config(mesh="simple", solver="standard", mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1,
                                                                                              False),
                         solver_config=ns_cfg)
    else:
        cfg = sim.Config(mesh="simple", solver="passive_transport",
                         mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                         solver_config=sim.PassiveTransportConfig(ns_cfg, sim.ADEConfig(list(phi_faces), False)))
    out = sim.setup(cfg)
    s, surf, _, oUd = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], tol, set_zeroP, "multigrid", list(final))
    return out, s, surf


def _lid(s, surf, U_d_api, gamma_d):
    """Lid u = 1 on the top face (cavity.py:115-117); gamma_d of the API equals the oracle's."""
    assert np.array_equal(np.asarray(gamma_d), s.gamma_d)
    return ns.cavity_lid(s, surf, np.zeros(len(s.gamma_d)))


def _state(coords, nn, gamma_d, U_d, seed=5):
    r = np.random.default_rng(seed)
    x, y = coords[:, 0], coords[:, 1]
    U = np.concatenate([np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y)])
    U = U + 0.05 * r.standard_normal(U.size)
    U[gamma_d] = U_d
    P = 0.1 * r.standard_normal(nn)
    phi = ((x > 0.2) & (x < 0.6) & (y > 0.3) & (y < 0.7)).astype(float)
    return U, P, phi


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_ns_step_vjp_matches_oracle(set_zeroP):
    n, dt, rho, mu = 32, 2e-3, 1.0, 1e-2
    print(f"\n[case] NS step VJP {n}x{n} set_zeroP={set_zeroP}", flush=True)
    (compute, timestep, coords, surf, h, nn, gamma_d, U_d), s, osurf = _setup(n, "standard", set_zeroP)
    U_d = _lid(s, osurf, U_d, gamma_d)
    U_n, P_n, _ = _state(coords, nn, gamma_d, U_d)
    gU, gP = RNG.standard_normal(2 * nn), RNG.standard_normal(nn)
    F_p, F = compute(U_d, None, 0, 0)
    ctx = timestep.context
    (gUn, gPn), (U, P), its = ctx.four_step_timestep_vjp(U_n, P_n, dt, rho, mu, gU, gP, U_d, F_p, F)
    oF_p = _np(F_p) if hasattr(F_p, "detach") else F_p
    ogUn, ogPn, fw, (m_it, p_it) = oadj.four_step_vjp(s, U_n, P_n, dt, rho, mu, U_d, oF_p, 0, gU, gP)
    print(f"  iters {its} oracle {(fw['U_it'], fw['P_it'], p_it, m_it)}  rel gU_n {_rel(_np(gUn), ogUn):.3e}"
          f"  gP_n {_rel(_np(gPn), ogPn):.3e}", flush=True)
    assert _rel(_np(U), fw["U"]) < 1e-10 and _rel(_np(P), fw["P"]) < 1e-10
    assert _rel(_np(gUn), ogUn) < 1e-9, _rel(_np(gUn), ogUn)
    assert _rel(_np(gPn), ogPn) < 1e-9, _rel(_np(gPn), ogPn)
    assert its[:3] == (fw["U_it"], fw["P_it"], p_it) and abs(its[3] - m_it) <= 1


@pytest.mark.parametrize("phi_faces,supg", [((), False), ((3,), False), ((), True), ((3,), True)])
def test_passive_step_vjp_matches_oracle(phi_faces, supg):
    n, dt, rho, mu, D = 32, 2e-3, 1.0, 1e-2, 1e-3
    print(f"\n[case] passive step VJP {n}x{n} phi_faces={phi_faces} supg={supg}", flush=True)
    (compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi), s, osurf = _setup(
        n, "passive", phi_faces=phi_faces)
    U_d = _lid(s, osurf, U_d, gamma_d)
    U_n, P_n, phi_n = _state(coords, nn, gamma_d, U_d)
    phi_d = None if gamma_d_phi is None else np.linspace(0.0, 1.0, len(gamma_d_phi))
    gU, gP, gphi = RNG.standard_normal(2 * nn), RNG.standard_normal(nn), RNG.standard_normal(nn)
    ctx = timestep.context
    F_p, F = compute(U_d, None, 0, 0)
    (gUn, gPn, gfn), (U, P, phi), its = ctx.passive_timestep_vjp(U_n, P_n, phi_n, dt, rho, mu, D, gU, gP, gphi,
                                                                 U_d, phi_d, F_p, F, supg)
    oF_p = _np(F_p)
    ogUn, ogPn, ogfn, (oU, oP, ophi), oits = oadj.passive_vjp(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, phi_d, oF_p,
                                                              0, gamma_d_phi, gU, gP, gphi, supg=supg)
    print(f"  iters {its[3:]} oracle {oits}  rel " + " ".join(
        f"{_rel(_np(a), b):.3e}" for a, b in ((gUn, ogUn), (gPn, ogPn), (gfn, ogfn))), flush=True)
    assert _rel(_np(phi), ophi) < 1e-10 and _rel(_np(U), oU) < 1e-10
    for a, b in ((gUn, ogUn), (gPn, ogPn), (gfn, ogfn)):
        assert _rel(_np(a), b) < 1e-9, _rel(_np(a), b)
    # The transposed BiCGSTAB solves start from the forward x0 against a random cotangent, so
    # their residual falls from |A^T x0| >> |b| and their last iterations run where rounding
    # dominates: the iteration at which it first drops below tol moves with the summation order
    # of every earlier solve of the step (GPU gathers and block reductions vs the oracle's
    # scatter), and the counts may differ by a few.  The cotangents above agree to 1e-9 and the
    # symmetric pressure CG count is exact.
    assert its[4] == oits[1] and abs(its[3] - oits[0]) <= 5 and abs(its[5] - oits[2]) <= 5


def _loss_setup(n, nsteps, tol=1e-10):
    (compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi), s, osurf = _setup(n, "passive",
                                                                                                      tol=tol)
    U_d = _lid(s, osurf, U_d, gamma_d)
    U0, P0, phi0 = _state(coords, nn, gamma_d, U_d)
    target = 0.1 * RNG.standard_normal(nn)
    return compute, timestep, s, nn, U_d, U0, P0, phi0, target


@pytest.mark.parametrize("supg", [False, True])
def test_rollout_gradient_matches_oracle(supg):
    import torch
    import jax_fem_cfd.setup.simulation as sim
    n, nsteps, dt, rho, mu, D = 16, 6, 2e-3, 1.0, 1e-2, 1e-3
    print(f"\n[case] rollout gradient {n}x{n} steps={nsteps} checkpoint=4 supg={supg}", flush=True)
    compute, timestep, s, nn, U_d, U0, P0, phi0, target = _loss_setup(n, nsteps)
    dev = torch.device("cuda", 0)
    U0t = torch.tensor(U0, dtype=torch.float64, device=dev, requires_grad=True)
    phi0t = torch.tensor(phi0, dtype=torch.float64, device=dev, requires_grad=True)
    P0t = torch.tensor(P0, dtype=torch.float64, device=dev)
    tt = torch.tensor(target, dtype=torch.float64, device=dev)
    if supg:
        from jax_fem_cfd.solvers import adjoint
        U, P, phi = adjoint.rollout(timestep.context, nsteps, U0t, P0t, phi0t, dt, rho, mu, D, U_d, supg=True,
                                   checkpoint=4)
    else:
        U, P, phi = sim.timestep_loop_scan(dt, nsteps, compute, timestep, rho, mu, D, U_d, None, U0t, P0t, phi0t,
                                           None, checkpoint=4)
    loss = 0.5 * torch.sum((phi - tt) ** 2)
    loss.backward()
    _, _, ophi = oadj.rollout(s, nsteps, U0, P0, phi0, dt, rho, mu, D, U_d, None, 0, 0, None, supg=supg)
    assert _rel(_np(phi), ophi) < 1e-10
    ogU, _, ogphi = oadj.rollout_vjp(s, nsteps, U0, P0, phi0, dt, rho, mu, D, U_d, None, 0, 0, None, None, None,
                                     ophi - target, checkpoint=4, supg=supg)
    assert _rel(_np(U0t.grad), ogU) < 1e-9, _rel(_np(U0t.grad), ogU)
    assert _rel(_np(phi0t.grad), ogphi) < 1e-9, _rel(_np(phi0t.grad), ogphi)


def test_rollout_gradient_matches_finite_differences_on_gpu():
    import torch
    import jax_fem_cfd.setup.simulation as sim
    n, nsteps, dt, rho, mu, D = 64, 20, 1e-3, 1.0, 1e-3, 1e-4
    print(f"\n[case] rollout gradient vs FD {n}x{n} steps={nsteps} checkpoint=5", flush=True)
    compute, timestep, s, nn, U_d, U0, P0, phi0, target = _loss_setup(n, nsteps, tol=1e-12)
    dev = torch.device("cuda", 0)
    P0t = torch.tensor(P0, dtype=torch.float64, device=dev)
    phi0t = torch.tensor(phi0, dtype=torch.float64, device=dev)
    tt = torch.tensor(target, dtype=torch.float64, device=dev)

    def loss_of(U):
        _, _, phi = sim.timestep_loop_scan(dt, nsteps, compute, timestep, rho, mu, D, U_d, None, U, P0t, phi0t,
                                           None, checkpoint=5)
        return 0.5 * torch.sum((phi - tt) ** 2)

    U0t = torch.tensor(U0, dtype=torch.float64, device=dev, requires_grad=True)
    loss_of(U0t).backward()
    g = U0t.grad.detach()
    with torch.no_grad():
        for seed in range(2):
            d = torch.tensor(np.random.default_rng(seed).standard_normal(2 * nn), dtype=torch.float64, device=dev)
            eps = 1e-3
            fd = (loss_of(U0t.detach() + eps * d) - loss_of(U0t.detach() - eps * d)).item() / (2 * eps)
            an = torch.dot(g, d).item()
            print(f"  FD {fd:.12e}  adjoint {an:.12e}", flush=True)
            assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)


def test_vjp_argument_errors():
    from jax_fem_cfd._native import NativeError
    (compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi), s, osurf = _setup(8, "passive")
    ctx = timestep.context
    U_n, P_n, phi_n = _state(coords, nn, gamma_d, U_d)
    import jax_fem_cfd._native as nat
    Ud, Pd, fd = nat.as_dev(U_n), nat.as_dev(P_n), nat.as_dev(phi_n)
    with pytest.raises(NativeError):  # forward outputs must not alias the inputs
        ctx.passive_timestep_vjp(Ud, Pd, fd, 1e-3, 1.0, 1e-2, 1e-3, gphi=np.ones(nn), U_d=U_d,
                                 out=(Ud, ctx._empty(nn), ctx._empty(nn), ctx._empty(2 * nn), ctx._empty(nn),
                                      ctx._empty(nn)))


def test_zero_cotangents_give_zero_gradients():
    (compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi), s, osurf = _setup(16, "passive")
    ctx = timestep.context
    U_n, P_n, phi_n = _state(coords, nn, gamma_d, U_d)
    z2, z1 = np.zeros(2 * nn), np.zeros(nn)
    (gUn, gPn, gfn), _, its = ctx.passive_timestep_vjp(U_n, P_n, phi_n, 1e-3, 1.0, 1e-2, 1e-3, z2, z1, z1, U_d)
    assert its[3:] == (0, 0, 0)
    assert not _np(gUn).any() and not _np(gPn).any() and not _np(gfn).any()
