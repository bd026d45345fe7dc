"""Hand-written reverse-mode products (VJPs) of the Navier-Stokes step and of the
passive-scalar step, and the checkpointed gradient of a rollout (oracle,
test-only).

The reference differentiates its rollout (jax_fem_cfd/setup/simulation.py:362-374,
``lax.scan`` over ``passive_transport.timestep``) with JAX reverse-mode AD.  Every
linear solve goes through ``lax.custom_linear_solve``
(jax_fem_cfd/utils/jax_iterative_solvers.py:201-211), so it is differentiated
implicitly:
  x = A(theta)^-1 b   ->   bbar = A^-T xbar,   thetabar = -<bbar, (dA/dtheta) x>,
with A^-T xbar solved by the same iterative method and the same closure values
(x0, tol, preconditioner) as the forward solve -- BiCGSTAB on the transposed
operator (``transpose_solve=isolve_solve``), and for CG (``symmetric=True``) the
forward operator itself.  x0 receives no cotangent.

The masked operators  A y = where(mask, y, Ahat y)  have the transpose
A^T y = Ahat^T (y masked to zero on the Dirichlet rows) + where(mask, y, 0).

Ce(U, beta) (element_calc.py:165-185) is linear in U, so
<w, C(U) x> = sum_{node I, comp d} U_Id g_Id(w, x) with
g_Id = sum_e sum_ab w_a T_d[I, a, b] x_b and T_d[I] = Ce(unit velocity at local
node I, component d).  T is built here from the oracle's own Ce routine.

The steps restated here are the NS step (standard.py:148-165) and the passive step
(passive_transport.py:64-97), plain and SUPG.  With SUPG the scalar matrices
A_e = (Me + tau Ce(U_n)^T)/dt + (Ce(U) + D Le + tau Ct(U_n))/2 and
B_e = (Me + tau Ce(U_n)^T)/dt - (Ce(U_n) + D Le + tau Ct(U_n))/2, tau = tau(U_n, Ce(U))
(element_calc.py:188-229) depend nonlinearly on U_n and U; the cotangents of
sum_e w_e^T (B_e phi_n - A_e phi) are taken here by complex-step differentiation of the
oracle's own (pinned) element routines, with a 3x3 node colouring so that no element
holds two perturbed DOFs -- exact to rounding, and independent of any hand derivation.
Pinned by: adjoint identities <A x, y> = <x, A^T y> for every transposed operator
and central finite differences of the forward map (tests/test_adjoint_oracle.py).
"""
import numpy as np

from . import elements as el
from . import emult
from .ns import four_step_timestep, update_phi  # noqa: F401  (forward pieces reused)
from .solvers import bicgstab, cg


# ------------------------------------------------------------------------------ building blocks
def conv_tensor(s, beta):
    """T[d, I, a, b] = d Ce[a, b] / d U_{I, d} for one element of the uniform grid (any element
    type: nen = nodes per element)."""
    nen = s.nconn.shape[1]
    nconn1 = np.arange(nen, dtype=np.int64)[None, :]
    T = np.zeros((s.ndim, nen, nen, nen))
    for d in range(s.ndim):
        for i in range(nen):
            U = np.zeros(nen * s.ndim)
            U[d * nen + i] = 1.0
            T[d, i] = el.conservative_convec_elem_calc(nconn1, U, s.wq, s.N, s.Nderiv, nen, s.ndim, beta)[0]
    return T


def conv_grad(s, T, w, x, ncomp):
    """g = d/dU <w, C(U) x> summed over ncomp components (w, x of length ncomp*nn)."""
    nn, nconn = s.nn, s.nconn
    out = np.zeros(s.ndim * nn)
    for c in range(ncomp):
        we = w[c * nn:(c + 1) * nn][nconn]
        xe = x[c * nn:(c + 1) * nn][nconn]
        contrib = np.einsum("diab,ea,eb->edi", T, we, xe)
        for d in range(s.ndim):
            out[d * nn:(d + 1) * nn] += np.bincount(nconn.ravel(), contrib[:, d, :].ravel(), minlength=nn)
    return out


def _elem_mult_T(x, A_all, nconn, nn):
    """sum_e A_e^T x_e (per-element matrices (ne, nen, nen))."""
    return np.bincount(nconn.ravel(), np.einsum("eba,eb->ea", A_all, x[nconn]).ravel(), minlength=nn)


def _solve_or_zero(method, A, b, **kw):
    """An exactly-zero right-hand side has the solution 0 (0 iterations).  The reference's
    solvers would iterate from x0 towards a zero residual norm instead (atol2 = 0)."""
    if float(b @ b) == 0.0:
        return np.zeros_like(b), 0
    return method(A, b, **kw)


def _mask(n, idx):
    m = np.zeros(n, dtype=bool)
    if idx is not None:
        m[np.asarray(idx)] = True
    return m


def momentum_ops(s, Ce, rho, mu, dt, theta=0.5):
    """(A, A^T) of the masked momentum operator (standard.py:20-28)."""
    nn, ndim, nconn = s.nn, s.ndim, s.nconn
    mU = _mask(ndim * nn, s.gamma_d)

    def Ahat(y, trans):
        out = (rho / dt) * emult.consistent_mass_mult(y, s.Me, nconn, nn, ndim)
        for c in range(ndim):
            yc = y[c * nn:(c + 1) * nn]
            cy = _elem_mult_T(yc, Ce, nconn, nn) if trans else emult.scalar_visc_conv_mult(yc, Ce, 0 * s.Le, 0.0,
                                                                                            nconn, nn)
            out[c * nn:(c + 1) * nn] += theta * (rho * cy + mu * emult.laplacian_element_mult(yc, s.Le, nconn, nn))
        return out

    def A(y):
        return np.where(mU, y, Ahat(y, False))

    def AT(y):
        return Ahat(np.where(mU, 0.0, y), True) + np.where(mU, y, 0.0)
    return A, AT


def scalar_ops(s, Ce, D, dt, gamma_d_phi):
    """(A, A^T) of the masked Crank-Nicolson scalar operator (passive_transport.py:41-48)."""
    nn, nconn = s.nn, s.nconn
    m = _mask(nn, gamma_d_phi)

    def Ahat(y, trans):
        cy = _elem_mult_T(y, Ce, nconn, nn) if trans else emult.scalar_visc_conv_mult(y, Ce, 0 * s.Le, 0.0, nconn,
                                                                                      nn)
        return ((1 / dt) * emult.consistent_mass_mult(y, s.Me, nconn, nn, 1)
                + 0.5 * (cy + D * emult.laplacian_element_mult(y, s.Le, nconn, nn)))

    def A(y):
        return np.where(m, y, Ahat(y, False))

    def AT(y):
        return Ahat(np.where(m, 0.0, y), True) + np.where(m, y, 0.0)
    return A, AT


# ------------------------------------------------------------------------------ forward with intermediates
def ns_forward(s, U_n, P_n, dt, rho, mu, U_d, F_p, F, theta=0.5):
    """four_step_timestep (standard.py:148-165) keeping U_hat, U_star and the solve x0s."""
    from .ns import solve_momentum, solve_pressure
    nn, ndim, nconn = s.nn, s.ndim, s.nconn
    Ce = el.conservative_convec_elem_calc(nconn, U_n, s.wq, s.N, s.Nderiv, nn, ndim, 0.5)
    U_hat, U_it = solve_momentum(U_n, P_n, s.Me, s.Le, s.Ge, Ce, rho, mu, dt, nconn, nn, ndim, s.gamma_d, U_d, F,
                                 s.tol, theta, 1)
    U_star = U_hat + (dt / rho) * emult.gradient_element_mult(P_n, s.Ge, nconn, nn, ndim) / s.M
    P, P_it = solve_pressure(U_star, P_n, s.Le, s.Ge, rho, dt, nconn, nn, ndim, s.gamma_p, F_p, s.set_zeroP,
                             s.M_op, s.tol, 0)
    U = U_star - (dt / rho) * emult.gradient_element_mult(P, s.Ge, nconn, nn, ndim) / s.M
    if s.gamma_d is not None:
        U = U.copy()
        U[s.gamma_d] = U_d
    return dict(Ce=Ce, U_hat=U_hat, U_star=U_star, P=P, U=U, U_it=U_it, P_it=P_it)


# ------------------------------------------------------------------------------ NS step VJP
def four_step_vjp(s, U_n, P_n, dt, rho, mu, U_d, F_p, F, gU, gP, theta=0.5, fw=None):
    """Cotangents (gU_n, gP_n) of the NS step for output cotangents (gU, gP).
    Returns (gU_n, gP_n, fw, (mom_adj_iters, p_adj_iters))."""
    nn, ndim, nconn = s.nn, s.ndim, s.nconn
    if fw is None:
        fw = ns_forward(s, U_n, P_n, dt, rho, mu, U_d, F_p, F, theta)
    mU = _mask(ndim * nn, s.gamma_d)
    G = lambda p: emult.gradient_element_mult(p, s.Ge, nconn, nn, ndim)  # noqa: E731
    Dv = lambda u: emult.divergence_element_mult(u, s.Ge, nconn, nn, ndim)  # noqa: E731
    gP = np.zeros(nn) if gP is None else gP

    # U = U* - (dt/rho) G P / M, then U[gamma_d] = U_d
    gUi = np.where(mU, 0.0, gU)
    gUs = gUi.copy()
    gPt = gP - (dt / rho) * Dv(gUi / s.M)
    # P = pin(P') or P' - mean(P')
    if s.set_zeroP:
        gPt = gPt.copy()
        gPt[s.gamma_p] = 0.0
        x0p = P_n.copy()
        x0p[s.gamma_p] = 0.0
    else:
        gPt = gPt - gPt.mean()
        x0p = P_n

    def lhs_p(P):
        res = emult.laplacian_element_mult(P, s.Le, nconn, nn)
        if s.set_zeroP:
            res[s.gamma_p] = 0
        return res

    # CG is marked symmetric: the transposed solve is the forward solve (same x0, M)
    mu_p, p_it = _solve_or_zero(cg, lhs_p, gPt, x0=x0p, tol=s.tol, maxiter=1000, M=s.M_op)
    # rhs = rho/dt (D U* - F_p), pinned rows zeroed / mean removed
    if s.set_zeroP:
        mu_p = mu_p.copy()
        mu_p[s.gamma_p] = 0.0
    else:
        mu_p = mu_p - mu_p.mean()
    gUs = gUs + (rho / dt) * G(mu_p)
    # U* = U_hat + (dt/rho) G P_n / M
    gPn = (dt / rho) * Dv(gUs / s.M)

    # U_hat = A_m(U_n)^-1 b_m(U_n, P_n): transposed BiCGSTAB from the forward x0
    Ce = fw["Ce"]
    _, lhs_mT = momentum_ops(s, Ce, rho, mu, dt, theta)

    x0m = U_n.copy()
    if s.gamma_d is not None:
        x0m[s.gamma_d] = U_d
    mu_m, m_it = _solve_or_zero(bicgstab, lhs_mT, gUs, x0=x0m, tol=s.tol)
    w = np.where(mU, 0.0, mu_m)
    # b_m = (rho/dt) M U_n - G P_n - (1-theta)(rho C(U_n) + mu L) U_n + F ; A_m depends on U_n via C
    gUn = (rho / dt) * emult.consistent_mass_mult(w, s.Me, nconn, nn, ndim)
    for c in range(ndim):
        wc = w[c * nn:(c + 1) * nn]
        gUn[c * nn:(c + 1) * nn] -= (1 - theta) * (rho * _elem_mult_T(wc, Ce, nconn, nn)
                                                   + mu * emult.laplacian_element_mult(wc, s.Le, nconn, nn))
    T = conv_tensor(s, 0.5)
    gUn -= rho * (theta * conv_grad(s, T, w, fw["U_hat"], ndim) + (1 - theta) * conv_grad(s, T, w, U_n, ndim))
    gPn = gPn - Dv(w)
    return gUn, gPn, fw, (m_it, p_it)


# ------------------------------------------------------------------------------ SUPG element matrices
def supg_matrices(s, U_n, U, D, dt):
    """Per-element SUPG scalar matrices (A_e, B_e) of passive_transport.py:83-91, from the
    oracle's element routines (complex-safe: einsum, sqrt of sums of squares, real powers)."""
    nconn, nn = s.nconn, s.nn
    Ce_n = el.conservative_convec_elem_calc(nconn, U_n, s.wq, s.N, s.Nderiv, nn, 2, 1)
    Ce = el.conservative_convec_elem_calc(nconn, U, s.wq, s.N, s.Nderiv, nn, 2, 1)
    Ct = el.Ctau_tilde_elem_calc(nconn, U_n, s.wq, s.N, s.Nderiv, nn, 2)
    tau = el.calculate_tau(nconn, U_n, Ce, Ct, s.N, nn, 2, D, dt, 2)
    Mt = s.Me[None] + tau[:, None, None] * Ce_n.transpose(0, 2, 1)
    K = D * s.Le[None] + tau[:, None, None] * Ct
    return Mt / dt + 0.5 * (Ce + K), Mt / dt - 0.5 * (Ce_n + K)


def supg_param_grad(s, U_n, U, w, phi, phi_n, D, dt, h=1e-30):
    """(d/dU_n, d/dU) of sum_e w_e^T (B_e phi_n,e - A_e phi_e), by complex step.  Nodes of one
    colour (i mod 3, j mod 3) share no element, so one complex evaluation per colour and
    component gives the derivative of every element term w.r.t. its single perturbed DOF."""
    nn, nconn = s.nn, s.nconn
    nnx = int(np.unique(s.coords[:, 0]).size)
    ii, jj = np.arange(nn) % nnx, np.arange(nn) // nnx
    we, pe, pne = w[nconn], phi[nconn], phi_n[nconn]

    def ell(Un, Uc):
        A, B = supg_matrices(s, Un, Uc, D, dt)
        return np.einsum("ea,eab,eb->e", we, B, pne) - np.einsum("ea,eab,eb->e", we, A, pe)

    out = []
    for which in (0, 1):
        g = np.zeros(2 * nn)
        for ci in range(3):
            for cj in range(3):
                sel = (ii % 3 == ci) & (jj % 3 == cj)
                for d in range(2):
                    P = np.zeros(2 * nn, dtype=complex)
                    P[d * nn:(d + 1) * nn][sel] = 1j * h
                    Un = U_n + (P if which == 0 else 0)
                    Uc = U + (P if which == 1 else 0)
                    de = ell(Un, Uc).imag / h                       # (ne,)
                    hit = sel[nconn]                                # (ne, 4): the element's perturbed node
                    node = np.where(hit.any(axis=1), nconn[np.arange(nconn.shape[0]), hit.argmax(axis=1)], -1)
                    ok = node >= 0
                    g[d * nn:(d + 1) * nn] += np.bincount(node[ok], de[ok], minlength=nn)
        out.append(g)
    return out[0], out[1]


# ------------------------------------------------------------------------------ passive step VJP
def passive_vjp(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi, gU, gP, gphi, supg=False):
    """Cotangents (gU_n, gP_n, gphi_n) of the passive step (passive_transport.py:64-97, plain
    or SUPG).  Returns (gU_n, gP_n, gphi_n, outputs (U, P, phi), adjoint iteration counts)."""
    from .ns import passive_timestep
    nn, ndim, nconn = s.nn, s.ndim, s.nconn
    fw = ns_forward(s, U_n, P_n, dt, rho, mu, U_d, F_p, F)
    U = fw["U"]
    _, _, phi, _, _, _ = passive_timestep(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi,
                                          supg)
    mphi = _mask(nn, gamma_d_phi)
    if supg:
        A_all, B_all = supg_matrices(s, U_n, U, D, dt)
    else:
        Ce_n = el.conservative_convec_elem_calc(nconn, U_n, s.wq, s.N, s.Nderiv, nn, ndim, 1)
        Ce = el.conservative_convec_elem_calc(nconn, U, s.wq, s.N, s.Nderiv, nn, ndim, 1)
        M_all = np.broadcast_to(s.Me, Ce.shape)
        K = D * s.Le[None]
        A_all, B_all = M_all / dt + 0.5 * (Ce + K), M_all / dt - 0.5 * (Ce_n + K)

    def lhs_sT(y):
        return _elem_mult_T(np.where(mphi, 0.0, y), A_all, nconn, nn) + np.where(mphi, y, 0.0)

    gU_tot = np.zeros(ndim * nn) if gU is None else gU
    gphin, gUn_s, s_it = np.zeros(nn), 0.0, 0
    if gphi is not None:
        x0s = phi_n.copy()
        if gamma_d_phi is not None:
            x0s[gamma_d_phi] = phi_d
        mu_s, s_it = _solve_or_zero(bicgstab, lhs_sT, gphi, x0=x0s, tol=s.tol)
        w = np.where(mphi, 0.0, mu_s)
        gphin = _elem_mult_T(w, B_all, nconn, nn)
        if supg:
            gUn_s, gUs = supg_param_grad(s, U_n, U, w, phi, phi_n, D, dt)
            gU_tot = gU_tot + gUs
        else:
            T1 = conv_tensor(s, 1.0)
            gU_tot = gU_tot - 0.5 * conv_grad(s, T1, w, phi, 1)
            gUn_s = -0.5 * conv_grad(s, T1, w, phi_n, 1)
    gUn, gPn, _, (m_it, p_it) = four_step_vjp(s, U_n, P_n, dt, rho, mu, U_d, F_p, F, gU_tot, gP, fw=fw)
    return gUn + gUn_s, gPn, gphin, (U, fw["P"], phi), (s_it, p_it, m_it)


# ------------------------------------------------------------------------------ rollout gradient
def rollout(s, nsteps, U0, P0, phi0, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi, supg=False):
    """timestep_loop_scan (simulation.py:362-374) on the passive step: final (U, P, phi)."""
    from .ns import passive_timestep
    U, P, phi = U0, P0, phi0
    for _ in range(nsteps):
        U, P, phi, _, _, _ = passive_timestep(s, U, P, phi, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi, supg)
    return U, P, phi


def rollout_vjp(s, nsteps, U0, P0, phi0, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi, gU, gP, gphi,
                checkpoint=20, supg=False):
    """Reverse sweep over the rollout with states stored every `checkpoint` steps and each
    segment recomputed forward before it is swept backward.  Returns (gU0, gP0, gphi0)."""
    from .ns import passive_timestep
    ckpt = {0: (U0, P0, phi0)}
    U, P, phi = U0, P0, phi0
    for k in range(nsteps):
        U, P, phi, _, _, _ = passive_timestep(s, U, P, phi, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi, supg)
        if (k + 1) % checkpoint == 0 and k + 1 < nsteps:
            ckpt[k + 1] = (U, P, phi)
    nn = s.nn
    gU = np.zeros(s.ndim * nn) if gU is None else gU
    gP = np.zeros(nn) if gP is None else gP
    gphi = np.zeros(nn) if gphi is None else gphi
    for seg0 in sorted(ckpt, reverse=True):
        seg1 = min(seg0 + checkpoint, nsteps)
        states = [ckpt[seg0]]
        for _ in range(seg0, seg1 - 1):
            Ua, Pa, pa = states[-1]
            Ub, Pb, pb, _, _, _ = passive_timestep(s, Ua, Pa, pa, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi,
                                                   supg)
            states.append((Ub, Pb, pb))
        for Ua, Pa, pa in reversed(states):
            gU, gP, gphi, _, _ = passive_vjp(s, Ua, Pa, pa, dt, rho, mu, D, U_d, phi_d, F_p, F, gamma_d_phi, gU, gP,
                                             gphi, supg)
    return gU, gP, gphi
