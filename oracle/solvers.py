"""Krylov solvers with the reference's exact control flow (oracle, test-only).

Follows jax_fem_cfd/utils/jax_iterative_solvers.py.  ``A`` and ``M`` are
callables on NumPy vectors.  Both return (x, iters) where iters counts loop
body executions; BiCGSTAB returns -10 / -11 on breakdown like the reference.
"""
import numpy as np


def _dot(x, y):
    # IEEE semantics (inf/nan on division by zero), as in the reference
    return np.float64(np.dot(x, y))


def cg(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """Preconditioned CG.  _cg_solve jax_iterative_solvers.py:87-120, wrapper
    cg :214-269 and _isolve :174-211 (x0 defaults to zeros, maxiter to
    10*size).  Stopping test on the recursive residual:
    r.r > max(tol^2 b.b, atol^2) and k < maxiter, where r.r is gamma = r.z
    when M is the identity and is recomputed otherwise (:95-98)."""
    b = np.asarray(b, dtype=float)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=float)
    if maxiter is None:
        maxiter = 10 * b.size
    ident = M is None
    if ident:
        M = lambda v: v  # noqa: E731
    atol2 = max(tol * tol * _dot(b, b), atol * atol)
    r = b - A(x)
    z = M(r)
    p = z
    gamma = _dot(r, z)
    k = 0
    while True:
        rs = gamma if ident else _dot(r, r)
        if not (rs > atol2 and k < maxiter):
            break
        Ap = A(p)
        alpha = gamma / _dot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = M(r)
        gamma_new = _dot(r, z)
        beta = gamma_new / gamma
        p = z + beta * p
        gamma = gamma_new
        k += 1
    return x, k


def bicgstab(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """BiCGSTAB.  _bicgstab_solve jax_iterative_solvers.py:123-167, wrapper
    bicgstab :685-742.  rho = alpha = omega = 1, r0 = b - A x0,
    rhat = p = q = r0; loop while r.r > atol2 and k < maxiter and k >= 0;
    early exit on s.s < atol2; k = -11 when omega or alpha is 0, -10 when
    rho is 0."""
    b = np.asarray(b, dtype=float)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=float)
    if maxiter is None:
        maxiter = 10 * b.size
    if M is None:
        M = lambda v: v  # noqa: E731
    atol2 = max(tol * tol * _dot(b, b), atol * atol)
    r = b - A(x)
    rhat = r.copy()
    p = r.copy()
    q = r.copy()
    alpha = omega = rho = np.float64(1.0)
    k = 0
    with np.errstate(divide="ignore", invalid="ignore"):
        while _dot(r, r) > atol2 and k < maxiter and k >= 0:
            rho_ = _dot(rhat, r)
            beta = rho_ / rho * alpha / omega
            p = r + beta * (p - omega * q)
            phat = M(p)
            q = A(phat)
            alpha = rho_ / _dot(rhat, q)
            s = r - alpha * q
            exit_early = _dot(s, s) < atol2
            shat = M(s)
            t = A(shat)
            omega = _dot(t, s) / _dot(t, t)
            if exit_early:
                x = x + alpha * phat
                r = s
            else:
                x = x + (alpha * phat + omega * shat)
                r = s - omega * t
            k_ = -11 if (omega == 0 or alpha == 0) else k + 1
            k = -10 if rho_ == 0 else k_
            rho = rho_
    return x, k
