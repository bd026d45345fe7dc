"""Pins of the oracle's hand-written reverse-mode products (oracle/adjoint.py).

The reference's own gradient is JAX reverse-mode AD (jax absent here), so the
oracle adjoint is pinned by what the mathematics fixes:
  * <w, C(delta) x> = <g(w, x), delta> exactly (Ce is linear in U), with C(delta)
    built by the oracle's pinned Ce routine;
  * <A x, y> = <x, A^T y> for the transposed masked momentum and scalar operators;
  * central finite differences of the forward step and of a checkpointed rollout,
    with the solves run to 1e-13 so that solver noise stays far below the FD error.
A dropped term (the lhs dependence of a solve on U_n, a mask, the mean / pin
projections, the Dirichlet overwrite) moves the directional derivative by O(1).
"""
import numpy as np
import pytest

from oracle import adjoint as adj
from oracle import elements as el
from oracle import emult
from oracle import ns

RNG = np.random.default_rng(7)


def _case(n=8, set_zeroP=True, tol=1e-13, dirichlet=(0, 1, 2, 3)):
    s, surf, _, U_d = ns.build([n, n], [1.0, 1.0], False, list(dirichlet), [], tol, set_zeroP, "multigrid", [2, 2])
    if s.gamma_d is not None:
        U_d = ns.cavity_lid(s, surf, U_d)
    return s, surf, U_d


def _state(s, U_d, seed=3):
    r = np.random.default_rng(seed)
    x, y = s.coords[:, 0], s.coords[:, 1]
    U = np.concatenate([np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y)])
    U = U + 0.05 * r.standard_normal(U.size)
    if s.gamma_d is not None:
        U[s.gamma_d] = U_d
    P = 0.1 * r.standard_normal(s.nn)
    phi = ((x > 0.2) & (x < 0.6) & (y > 0.3) & (y < 0.7)).astype(float) + 0.01 * r.standard_normal(s.nn)
    return U, P, phi


def test_conv_grad_is_the_gradient_of_the_bilinear_form():
    s, _, _ = _case(6)
    nn = s.nn
    for beta in (0.5, 1.0):
        T = adj.conv_tensor(s, beta)
        for ncomp in (1, 2):
            w, x = RNG.standard_normal(ncomp * nn), RNG.standard_normal(ncomp * nn)
            delta = RNG.standard_normal(2 * nn)
            Ce = el.conservative_convec_elem_calc(s.nconn, delta, s.wq, s.N, s.Nderiv, nn, 2, beta)
            lhs = sum(w[c * nn:(c + 1) * nn] @ emult.scalar_visc_conv_mult(x[c * nn:(c + 1) * nn], Ce, 0 * s.Le, 0.0,
                                                                             s.nconn, nn) for c in range(ncomp))
            rhs = adj.conv_grad(s, T, w, x, ncomp) @ delta
            assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0)


def test_transposed_operators_satisfy_the_adjoint_identity():
    s, _, _ = _case(6)
    nn = s.nn
    U = RNG.standard_normal(2 * nn)
    Ce = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, nn, 2, 0.5)
    A, AT = adj.momentum_ops(s, Ce, 1.0, 1e-2, 1e-2)
    x, y = RNG.standard_normal(2 * nn), RNG.standard_normal(2 * nn)
    assert abs(A(x) @ y - x @ AT(y)) <= 1e-12 * abs(A(x) @ y)
    # the mask really is transposed: A^T is not A with Ce^T alone
    Ce1 = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, nn, 2, 1.0)
    for gd in (None, np.arange(0, nn, 5)):
        A, AT = adj.scalar_ops(s, Ce1, 1e-3, 1e-2, gd)
        x, y = RNG.standard_normal(nn), RNG.standard_normal(nn)
        assert abs(A(x) @ y - x @ AT(y)) <= 1e-12 * abs(A(x) @ y)


def _fd(f, xs, ds, eps):
    plus = f(*[x + eps * d for x, d in zip(xs, ds)])
    minus = f(*[x - eps * d for x, d in zip(xs, ds)])
    return (plus - minus) / (2 * eps)


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_ns_step_vjp_matches_finite_differences(set_zeroP):
    s, _, U_d = _case(8, set_zeroP=set_zeroP)
    U_n, P_n, _ = _state(s, U_d)
    nn = s.nn
    dt, rho, mu = 5e-3, 1.0, 1e-2
    F_p = el.u_dot_n_boundary_integral_1d(s.bconn, s.nvec, s.gamma_d, U_d, nn, s.coords)
    a, b = RNG.standard_normal(2 * nn), RNG.standard_normal(nn)

    def loss(U0, P0):
        U, P, _, _ = ns.four_step_timestep(s, U0, P0, dt, rho, mu, U_d, F_p, 0)
        return a @ U + b @ P

    gUn, gPn, _, _ = adj.four_step_vjp(s, U_n, P_n, dt, rho, mu, U_d, F_p, 0, a, b)
    for _ in range(3):
        dU, dP = RNG.standard_normal(2 * nn), RNG.standard_normal(nn)
        fd = _fd(loss, [U_n, P_n], [dU, dP], 1e-4)
        an = gUn @ dU + gPn @ dP
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    # the U_n cotangent alone (the Ce(U_n) terms included)
    dU = RNG.standard_normal(2 * nn)
    fd = _fd(lambda U0: loss(U0, P_n), [U_n], [dU], 1e-4)
    assert abs(fd - gUn @ dU) <= 1e-6 * abs(fd)


def test_passive_step_vjp_matches_finite_differences():
    s, _, U_d = _case(8)
    U_n, P_n, phi_n = _state(s, U_d)
    nn = s.nn
    dt, rho, mu, D = 5e-3, 1.0, 1e-2, 1e-3
    a, b, c = RNG.standard_normal(2 * nn), RNG.standard_normal(nn), RNG.standard_normal(nn)

    def loss(U0, P0, f0):
        U, P, phi, _, _, _ = ns.passive_timestep(s, U0, P0, f0, dt, rho, mu, D, U_d, None, 0, 0, None, False)
        return a @ U + b @ P + c @ phi

    gUn, gPn, gfn, outs, _ = adj.passive_vjp(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, None, 0, 0, None, a, b, c)
    U, P, phi, _, _, _ = ns.passive_timestep(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, None, 0, 0, None, False)
    assert np.array_equal(outs[0], U) and np.array_equal(outs[2], phi)
    for _ in range(3):
        ds = [RNG.standard_normal(2 * nn), RNG.standard_normal(nn), RNG.standard_normal(nn)]
        fd = _fd(loss, [U_n, P_n, phi_n], ds, 1e-4)
        an = gUn @ ds[0] + gPn @ ds[1] + gfn @ ds[2]
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_passive_step_vjp_with_scalar_dirichlet_nodes():
    s, surf, U_d = _case(8)
    U_n, P_n, phi_n = _state(s, U_d)
    nn = s.nn
    gdp = np.unique(np.concatenate([surf[3]]))
    phi_d = np.linspace(0.0, 1.0, gdp.size)
    dt, rho, mu, D = 5e-3, 1.0, 1e-2, 1e-3
    c = RNG.standard_normal(nn)

    def loss(U0, f0):
        _, _, phi, _, _, _ = ns.passive_timestep(s, U0, P_n, f0, dt, rho, mu, D, U_d, phi_d, 0, 0, gdp, False)
        return c @ phi

    gUn, _, gfn, _, _ = adj.passive_vjp(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, phi_d, 0, 0, gdp, None, None, c)
    ds = [RNG.standard_normal(2 * nn), RNG.standard_normal(nn)]
    fd = _fd(loss, [U_n, phi_n], ds, 1e-4)
    an = gUn @ ds[0] + gfn @ ds[1]
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_rollout_vjp_with_checkpoints_matches_finite_differences():
    s, _, U_d = _case(8, tol=1e-13)
    U0, P0, phi0 = _state(s, U_d)
    nn = s.nn
    dt, rho, mu, D, nsteps = 5e-3, 1.0, 1e-2, 1e-3, 5
    target = RNG.standard_normal(nn) * 0.1

    def loss(U):
        _, _, phi = adj.rollout(s, nsteps, U, P0, phi0, dt, rho, mu, D, U_d, None, 0, 0, None)
        return 0.5 * np.sum((phi - target) ** 2)

    _, _, phiN = adj.rollout(s, nsteps, U0, P0, phi0, dt, rho, mu, D, U_d, None, 0, 0, None)
    g2, _, _ = adj.rollout_vjp(s, nsteps, U0, P0, phi0, dt, rho, mu, D, U_d, None, 0, 0, None, None, None,
                               phiN - target, checkpoint=2)
    g5, _, _ = adj.rollout_vjp(s, nsteps, U0, P0, phi0, dt, rho, mu, D, U_d, None, 0, 0, None, None, None,
                               phiN - target, checkpoint=5)
    assert np.array_equal(g2, g5)  # checkpointing changes nothing: the recomputation is exact
    dU = RNG.standard_normal(2 * nn)
    fd = _fd(loss, [U0], [dU], 1e-4)
    assert abs(fd - g2 @ dU) <= 1e-6 * abs(fd), (fd, g2 @ dU)


def test_zero_cotangent_gives_zero_without_iterating():
    s, _, U_d = _case(6)
    U_n, P_n, phi_n = _state(s, U_d)
    z2, z1 = np.zeros(2 * s.nn), np.zeros(s.nn)
    gUn, gPn, gfn, _, its = adj.passive_vjp(s, U_n, P_n, phi_n, 5e-3, 1.0, 1e-2, 1e-3, U_d, None, 0, 0, None, z2,
                                            z1, z1)
    assert not gUn.any() and not gPn.any() and not gfn.any() and its == (0, 0, 0)


def test_supg_matrices_match_the_forward_step_operators():
    """The SUPG element matrices used by the adjoint are the operators of update_phi_SUPG
    (passive_transport.py:14-37): A phi = lhs(phi), B phi_n = rhs."""
    s, _, U_d = _case(8)
    U_n, P_n, phi_n = _state(s, U_d)
    U = U_n + 0.01 * RNG.standard_normal(U_n.size)
    D, dt = 1e-3, 5e-3
    A, B = adj.supg_matrices(s, U_n, U, D, dt)
    Ce_n = el.conservative_convec_elem_calc(s.nconn, U_n, s.wq, s.N, s.Nderiv, s.nn, 2, 1)
    Ce = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 2, 1)
    Ct = el.Ctau_tilde_elem_calc(s.nconn, U_n, s.wq, s.N, s.Nderiv, s.nn, 2)
    tau = el.calculate_tau(s.nconn, U_n, Ce, Ct, s.N, s.nn, 2, D, dt, 2)
    Mtau, Ctau = tau[:, None, None] * Ce_n.transpose(0, 2, 1), tau[:, None, None] * Ct
    x = RNG.standard_normal(s.nn)
    lhs = (1 / dt) * emult.scalar_consistent_mass_SUPG_mult(x, Mtau, s.Me, s.nconn, s.nn) + \
        0.5 * emult.scalar_stabilized_visc_conv_mult(x, Ce, Ctau, s.Le, D, s.nconn, s.nn)
    rhs = (1 / dt) * emult.scalar_consistent_mass_SUPG_mult(x, Mtau, s.Me, s.nconn, s.nn) - \
        0.5 * emult.scalar_stabilized_visc_conv_mult(x, Ce_n, Ctau, s.Le, D, s.nconn, s.nn)
    Ax = np.bincount(s.nconn.ravel(), np.einsum("eab,eb->ea", A, x[s.nconn]).ravel(), minlength=s.nn)
    Bx = np.bincount(s.nconn.ravel(), np.einsum("eab,eb->ea", B, x[s.nconn]).ravel(), minlength=s.nn)
    assert np.abs(Ax - lhs).max() <= 1e-12 * np.abs(lhs).max()
    assert np.abs(Bx - rhs).max() <= 1e-12 * np.abs(rhs).max()


def test_supg_passive_step_vjp_matches_finite_differences():
    s, _, U_d = _case(8)
    U_n, P_n, phi_n = _state(s, U_d)
    nn = s.nn
    dt, rho, mu, D = 5e-3, 1.0, 1e-2, 1e-3
    a, b, c = RNG.standard_normal(2 * nn), RNG.standard_normal(nn), RNG.standard_normal(nn)

    def loss(U0, P0, f0):
        U, P, phi, _, _, _ = ns.passive_timestep(s, U0, P0, f0, dt, rho, mu, D, U_d, None, 0, 0, None, True)
        return a @ U + b @ P + c @ phi

    gUn, gPn, gfn, _, _ = adj.passive_vjp(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, None, 0, 0, None, a, b, c,
                                          supg=True)
    for _ in range(3):
        ds = [RNG.standard_normal(2 * nn), RNG.standard_normal(nn), RNG.standard_normal(nn)]
        fd = _fd(loss, [U_n, P_n, phi_n], ds, 1e-4)
        an = gUn @ ds[0] + gPn @ ds[1] + gfn @ ds[2]
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    # the U_n and U (through tau, Ct, Ce) cotangents alone: phi cotangent only
    gUn, _, _, _, _ = adj.passive_vjp(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, None, 0, 0, None, None, None, c,
                                      supg=True)
    dU = RNG.standard_normal(2 * nn)
    fd = _fd(lambda U0: c @ ns.passive_timestep(s, U0, P_n, phi_n, dt, rho, mu, D, U_d, None, 0, 0, None, True)[2],
             [U_n], [dU], 1e-4)
    assert abs(fd - gUn @ dU) <= 1e-6 * abs(fd), (fd, gUn @ dU)


# ----------------------------------------------------------------------------- quad9
def _case_q2(n=4, set_zeroP=True, tol=1e-13):
    s, surf, _, U_d = ns.build([n, n], [1.0, 1.0], False, [0, 1, 2, 3], [], tol, set_zeroP, "multigrid", [2, 2],
                               shape_func_ord=2)
    return s, surf, ns.cavity_lid(s, surf, U_d)


def test_quad9_conv_grad_and_transposed_operators():
    """The same identities on quad9: <w, C(delta) x> = <g(w, x), delta> and <A x, y> = <x, A^T y>."""
    s, _, _ = _case_q2(3)
    nn = s.nn
    T = adj.conv_tensor(s, 0.5)
    w, x = RNG.standard_normal(2 * nn), RNG.standard_normal(2 * nn)
    delta = RNG.standard_normal(2 * nn)
    Ce = el.conservative_convec_elem_calc(s.nconn, delta, s.wq, s.N, s.Nderiv, nn, 2, 0.5)
    lhs = sum(w[c * nn:(c + 1) * nn] @ emult.scalar_visc_conv_mult(x[c * nn:(c + 1) * nn], Ce, 0 * s.Le, 0.0,
                                                                     s.nconn, nn) for c in range(2))
    assert abs(lhs - adj.conv_grad(s, T, w, x, 2) @ delta) <= 1e-12 * (abs(lhs) + 1.0)
    U = RNG.standard_normal(2 * nn)
    Ce = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, nn, 2, 0.5)
    A, AT = adj.momentum_ops(s, Ce, 1.0, 1e-2, 1e-2)
    x, y = RNG.standard_normal(2 * nn), RNG.standard_normal(2 * nn)
    assert abs(A(x) @ y - x @ AT(y)) <= 1e-12 * abs(A(x) @ y)


@pytest.mark.parametrize("set_zeroP", [True, False])
def test_quad9_ns_step_vjp_matches_finite_differences(set_zeroP):
    s, _, U_d = _case_q2(4, set_zeroP=set_zeroP)
    U_n, P_n, _ = _state(s, U_d)
    nn = s.nn
    dt, rho, mu = 5e-3, 1.0, 1e-2
    F_p = el.u_dot_n_boundary_integral_1d(s.bconn, s.nvec, s.gamma_d, U_d, nn, s.coords, el.lin3_shape_functions)
    a, b = RNG.standard_normal(2 * nn), RNG.standard_normal(nn)

    def loss(U0, P0):
        U, P, _, _ = ns.four_step_timestep(s, U0, P0, dt, rho, mu, U_d, F_p, 0)
        return a @ U + b @ P

    gUn, gPn, _, _ = adj.four_step_vjp(s, U_n, P_n, dt, rho, mu, U_d, F_p, 0, a, b)
    for _ in range(2):
        dU, dP = RNG.standard_normal(2 * nn), RNG.standard_normal(nn)
        fd = _fd(loss, [U_n, P_n], [dU, dP], 1e-4)
        an = gUn @ dU + gPn @ dP
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


@pytest.mark.parametrize("supg", [False, True])
def test_quad9_passive_step_vjp_matches_finite_differences(supg):
    """quad9 passive step, plain and SUPG: the oracle VJP's directional derivatives against central
    finite differences of the oracle's own step.  With SUPG the element-parameter cotangents come
    from complex steps over a node colouring (i mod 3, j mod 3): three consecutive node indices are
    distinct mod 3, so a quad9 element holds at most one perturbed node of a colour."""
    s, surf, U_d = _case_q2(4)
    s_p, *_ = ns.build([4, 4], [1.0, 1.0], False, [0, 1, 2, 3], [], 1e-13, True, "multigrid", [2, 2],
                       shape_func_ord=2, passive=True)
    s = s_p
    U_n, P_n, phi_n = _state(s, U_d)
    nn = s.nn
    dt, rho, mu, D = 5e-3, 1.0, 1e-2, 1e-3
    a, b, c = RNG.standard_normal(2 * nn), RNG.standard_normal(nn), RNG.standard_normal(nn)

    def loss(U0, P0, f0):
        U, P, phi, _, _, _ = ns.passive_timestep(s, U0, P0, f0, dt, rho, mu, D, U_d, None, 0, 0, None, supg)
        return a @ U + b @ P + c @ phi

    gUn, gPn, gfn, outs, _ = adj.passive_vjp(s, U_n, P_n, phi_n, dt, rho, mu, D, U_d, None, 0, 0, None, a, b, c,
                                             supg=supg)
    for _ in range(2):
        ds = [RNG.standard_normal(2 * nn), RNG.standard_normal(nn), RNG.standard_normal(nn)]
        fd = _fd(loss, [U_n, P_n, phi_n], ds, 1e-4)
        an = gUn @ ds[0] + gPn @ ds[1] + gfn @ ds[2]
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
