"""CG and BiCGSTAB with iteration counts, on GPU vectors.

Same call surface as jax_fem_cfd/utils/jax_iterative_solvers.py
(``cg(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None)`` and
``bicgstab(...)``, returning ``(x, iters)``) for arbitrary operator callables:
the recurrences run on the host, every vector operation (dot products,
axpys) runs in the HIP BLAS-1 kernels of libfcfd with deterministic
reductions.  The timestep itself does not use this path: its solves run as
fused device-side drivers (fcfd_solve_momentum / fcfd_pressure_cg).

Breakdown codes as in the reference: BiCGSTAB returns -10 when rhat.r = 0
and -11 when omega or alpha is 0 (:156-157).
"""
import numpy as np

from .. import _native as nat
from .._context import Context

_blas = {}


def _ctx():
    c = _blas.get("ctx")
    if c is None:
        c = Context(1, 1, 1.0, 1.0, False)
        _blas["ctx"] = c
    return c


def _dot(x, y):
    return _ctx().dot(x, y)


def _axpby(a, x, b, y):
    return _ctx().axpby(a, x, b, y)


def _apply(f, x):
    return nat.as_dev(f(x))


def cg(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """_cg_solve (jax_iterative_solvers.py:87-120): stop when
    r.r <= max(tol^2 b.b, atol^2) or k = maxiter; r.r is gamma = r.z when M is
    the identity and is recomputed otherwise."""
    t = nat.torch()
    b = nat.as_dev(b)
    x = t.zeros_like(b) if x0 is None else nat.as_dev(x0).clone()
    if maxiter is None:
        maxiter = 10 * b.numel()
    ident = M is None
    atol2 = max(tol * tol * _dot(b, b), atol * atol)
    r = b.clone()
    _axpby(-1.0, _apply(A, x), 1.0, r)
    z = r.clone() if ident else _apply(M, r)
    p = z.clone()
    gamma = _dot(r, z)
    k = 0
    while True:
        rs = gamma if ident else _dot(r, r)
        if not (rs > atol2 and k < maxiter):
            break
        Ap = _apply(A, p)
        alpha = gamma / _dot(p, Ap)
        _axpby(alpha, p, 1.0, x)
        _axpby(-alpha, Ap, 1.0, r)
        z = r.clone() if ident else _apply(M, r)
        gamma_new = _dot(r, z)
        beta = gamma_new / gamma
        _axpby(1.0, z, beta, p)
        gamma = gamma_new
        k += 1
    return x, k


def _div(a, b):
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float64(a) / np.float64(b))


def bicgstab(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """_bicgstab_solve (jax_iterative_solvers.py:123-167)."""
    t = nat.torch()
    b = nat.as_dev(b)
    x = t.zeros_like(b) if x0 is None else nat.as_dev(x0).clone()
    if maxiter is None:
        maxiter = 10 * b.numel()
    Mf = (lambda v: v) if M is None else (lambda v: _apply(M, v))
    atol2 = max(tol * tol * _dot(b, b), atol * atol)
    r = b.clone()
    _axpby(-1.0, _apply(A, x), 1.0, r)
    rhat, p, q = r.clone(), r.clone(), r.clone()
    alpha = omega = rho = 1.0
    k = 0
    while _dot(r, r) > atol2 and k < maxiter and k >= 0:
        rho_ = _dot(rhat, r)
        beta = _div(rho_, rho) * alpha
        beta = _div(beta, omega)
        _axpby(-omega, q, 1.0, p)           # p - omega q
        _axpby(1.0, r, beta, p)             # r + beta (p - omega q)
        phat = Mf(p)
        q = _apply(A, phat)
        alpha = _div(rho_, _dot(rhat, q))
        s = r.clone()
        _axpby(-alpha, q, 1.0, s)
        exit_early = _dot(s, s) < atol2
        shat = Mf(s)
        tv = _apply(A, shat)
        omega = _div(_dot(tv, s), _dot(tv, tv))
        if exit_early:
            _axpby(alpha, phat, 1.0, x)
            r = s
        else:
            upd = shat.clone()
            _axpby(alpha, phat, omega, upd)  # alpha phat + omega shat
            _axpby(1.0, upd, 1.0, x)
            r = s.clone()
            _axpby(-omega, tv, 1.0, r)
        k_ = -11 if (omega == 0 or alpha == 0) else k + 1
        k = -10 if rho_ == 0 else k_
        rho = rho_
    return x, k
