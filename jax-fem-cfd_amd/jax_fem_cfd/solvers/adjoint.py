"""Differentiable rollouts: reverse mode by hand-written adjoint steps with
checkpointed recomputation.

The reference differentiates ``timestep_loop_scan`` (jax_fem_cfd/setup/simulation.py:362-374)
with JAX reverse-mode AD through ``lax.scan`` -- full storage of every step's
intermediates -- and differentiates each linear solve implicitly through
``lax.custom_linear_solve`` (jax_fem_cfd/utils/jax_iterative_solvers.py:201-211).

Here the rollout is one ``torch.autograd.Function``:
  forward   runs the steps on the GPU and keeps the state (U, P, phi) every
            ``checkpoint`` steps;
  backward  walks the segments last to first: recomputes the segment's states from
            its checkpoint, then calls fcfd_passive_timestep_vjp (or
            fcfd_four_step_timestep_vjp) once per step, last step first.  Each call
            re-runs its step to rebuild the intermediates (U_hat, Ce(U_n), the scalar
            element matrices) and then runs the adjoint solves: transposed BiCGSTAB for
            the scalar and momentum systems, MG-PCG for the pressure system.
Memory: one state per checkpoint plus one segment of states; no per-iteration tape.
"""
from .. import _native as nat


def _torch():
    return nat.torch()


class _Rollout:
    """Bound step / VJP calls of one context and one set of step parameters."""

    def __init__(self, ctx, dt, rho, mu, D, U_d, phi_d, F_p, F, supg, passive):
        self.ctx, self.dt, self.rho, self.mu, self.D = ctx, dt, rho, mu, D
        self.U_d, self.phi_d, self.F_p, self.F = U_d, phi_d, F_p, F
        self.supg, self.passive = bool(supg), passive

    def step(self, U, P, phi):
        c = self.ctx
        if self.passive:
            U, P, phi, *_ = c.passive_timestep(U, P, phi, self.dt, self.rho, self.mu, self.D, self.U_d, self.phi_d,
                                               self.F_p, self.F, self.supg)
            return U, P, phi
        U, P, *_ = c.four_step_timestep(U, P, self.dt, self.rho, self.mu, self.U_d, self.F_p, self.F)
        return U, P, None

    def vjp(self, U, P, phi, gU, gP, gphi):
        c = self.ctx
        if self.passive:
            (gUn, gPn, gfn), _, _ = c.passive_timestep_vjp(U, P, phi, self.dt, self.rho, self.mu, self.D, gU, gP,
                                                          gphi, self.U_d, self.phi_d, self.F_p, self.F, self.supg)
            return gUn, gPn, gfn
        (gUn, gPn), _, _ = c.four_step_timestep_vjp(U, P, self.dt, self.rho, self.mu, gU, gP, self.U_d, self.F_p,
                                                    self.F)
        return gUn, gPn, None


def _run(ro, nsteps, checkpoint, U, P, phi, keep):
    ckpts = {}
    for k in range(nsteps):
        if keep and k % checkpoint == 0:
            ckpts[k] = (U, P, phi)
        U, P, phi = ro.step(U, P, phi)
    return U, P, phi, ckpts


def _make_function():
    torch = _torch()

    class RolloutFunction(torch.autograd.Function):
        @staticmethod
        def forward(fctx, U0, P0, phi0, ro, nsteps, checkpoint):
            U0, P0 = nat.as_dev(U0.detach()), nat.as_dev(P0.detach())
            phi0 = nat.as_dev(phi0.detach()) if phi0 is not None else None
            U, P, phi, ckpts = _run(ro, nsteps, checkpoint, U0, P0, phi0, True)
            fctx.ro, fctx.nsteps, fctx.checkpoint, fctx.ckpts = ro, nsteps, checkpoint, ckpts
            fctx.has_phi = phi0 is not None
            if phi is None:
                phi = torch.zeros(0, dtype=torch.float64, device=U.device)
            return U, P, phi

        @staticmethod
        def backward(fctx, gU, gP, gphi):
            ro, nsteps, cp = fctx.ro, fctx.nsteps, fctx.checkpoint
            gU = None if gU is None else gU.contiguous()
            gP = None if gP is None else gP.contiguous()
            gphi = None if (gphi is None or not fctx.has_phi) else gphi.contiguous()
            for seg0 in sorted(fctx.ckpts, reverse=True):
                seg1 = min(seg0 + cp, nsteps)
                states = [fctx.ckpts[seg0]]
                for _ in range(seg0, seg1 - 1):
                    states.append(ro.step(*states[-1]))
                for U, P, phi in reversed(states):
                    gU, gP, gphi = ro.vjp(U, P, phi, gU, gP, gphi)
            fctx.ckpts = None
            return gU, gP, gphi, None, None, None

    return RolloutFunction


_FN = None


def rollout(ctx, nsteps, U0, P0, phi0, dt, rho, mu, D=0.0, U_d=None, phi_d=None, F_p=None, F=None, supg=False,
            checkpoint=20):
    """Run `nsteps` steps from (U0, P0, phi0) on a native context and return the final
    (U, P, phi) (phi None for the NS-only step when phi0 is None).  Differentiable w.r.t.
    U0, P0, phi0 through torch autograd: backward = checkpointed adjoint sweep."""
    global _FN
    torch = _torch()
    passive = phi0 is not None
    ro = _Rollout(ctx, dt, rho, mu, D, U_d, phi_d, F_p, F, supg, passive)
    tens = [x if isinstance(x, torch.Tensor) else nat.as_dev(x) for x in (U0, P0)]
    phi0 = (phi0 if isinstance(phi0, torch.Tensor) else nat.as_dev(phi0)) if passive else None
    needs = any(x.requires_grad for x in tens) or (passive and phi0.requires_grad)
    if not (needs and torch.is_grad_enabled()):
        U, P, phi, _ = _run(ro, nsteps, checkpoint, nat.as_dev(tens[0]), nat.as_dev(tens[1]),
                            nat.as_dev(phi0) if passive else None, False)
        return U, P, phi
    if _FN is None:
        _FN = _make_function()
    U, P, phi = _FN.apply(tens[0], tens[1], phi0, ro, int(nsteps), int(checkpoint))
    return U, P, (phi if passive else None)
