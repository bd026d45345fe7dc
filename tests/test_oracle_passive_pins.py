"""Independent pins of the oracle's passive-scalar step, SUPG stabilisation and pure-Neumann
V-cycle (jax_fem_cfd/solvers/passive_transport.py:14-97, core/element_calc.py:188-229,
core/multigrid.py:231-330).

What fixes the expected values here, none of it taken from the oracle:
  * closed-form Q1 element matrices on an hx x hy rectangle from 1D linear-element integrals
    (mass h/6 [[2,1],[1,2]], stiffness 1/h [[1,-1],[-1,1]], first-derivative [[-1/2,1/2],[-1/2,1/2]])
    in tensor product, assembled densely and solved with numpy.linalg.solve;
  * hand values of tau (1/3 for u = (1,0), h = 1, dt = 1, D + 1e-6 = 1/4) and its dt -> 0 limit;
  * invariants: with beta = 1 and walls, 1^T C = 0, so the CN step conserves 1^T M_c phi; with
    U = 0, a nodal cosine mode is a discrete eigenvector of (L, M_c) with the Q1 eigenvalue
    (6/h^2)(1 - cos t)/(2 + cos t) per direction and decays by the CN factor;
  * the pure-Neumann V-cycle restated with dense matrices, tensor-product linear interpolation
    and explicit mean projections.
Each group ends with planted mistakes (source-edited copies of the oracle function, here on
the CPU) that its check must reject: tau from U instead of U_n, Ct from U, M_tau without the
transpose, beta = 1/2 in the scalar Ce, r = 1, no D + 1e-6, and each mean projection of the
V-cycle whose removal can change its output (the other four are shown to be inert).
"""
import inspect
import textwrap

import numpy as np
import pytest

from oracle import elements as el
from oracle import emult
from oracle import mesh as ms
from oracle import multigrid as mg
from oracle import ns

RNG = np.random.default_rng(2024)
LOCAL_IJ = [(0, 0), (1, 0), (1, 1), (0, 1)]  # counter-clockwise quad4 corners [BL, BR, TR, TL]


# ----------------------------------------------------------------------------- helpers
def _mutant(func, old, new, count=1):
    """func with `old` replaced by `new` in its source, compiled in func's module namespace."""
    src = textwrap.dedent(inspect.getsource(func))
    assert src.count(old) == count, (old, src.count(old))
    g = dict(func.__globals__)
    exec(compile(src.replace(old, new), f"<mutant of {func.__name__}>", "exec"), g)
    return g[func.__name__]


def _line_1d(h):
    m = h / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])
    k = 1.0 / h * np.array([[1.0, -1.0], [-1.0, 1.0]])
    g = np.array([[-0.5, 0.5], [-0.5, 0.5]])  # g[i][k] = int X_i X_k'
    return m, k, g


def _q1_closed(hx, hy):
    """Exact Q1 element integrals on an hx x hy rectangle (tensor products of 1D integrals)."""
    mx, kx, gx = _line_1d(hx)
    my, ky, gy = _line_1d(hy)

    def tp(A, B):
        return np.array([[A[ia][ib] * B[ja][jb] for (ib, jb) in LOCAL_IJ] for (ia, ja) in LOCAL_IJ])

    out = dict(Me=tp(mx, my), Kxx=tp(kx, my), Kyy=tp(mx, ky), Cx=tp(gx, my), Cy=tp(mx, gy),
               Kxy=tp(gx.T, gy))  # Kxy[a][b] = int dN_a/dx dN_b/dy
    out["Le"] = out["Kxx"] + out["Kyy"]
    return out


def _ct_uniform(E, a, b):
    return a * a * E["Kxx"] + a * b * (E["Kxy"] + E["Kxy"].T) + b * b * E["Kyy"]


def _ce_uniform(E, a, b):
    return a * E["Cx"] + b * E["Cy"]  # beta irrelevant: a uniform field has zero divergence


def _assemble(nconn, nn, Ae):
    A = np.zeros((nn, nn))
    Ae = np.broadcast_to(Ae, (nconn.shape[0], 4, 4))
    for e, c in enumerate(nconn):
        A[np.ix_(c, c)] += Ae[e]
    return A


def _grid(nex, ney, lx, ly, tol=1e-13):
    s, surf, *_ = ns.build([nex, ney], [lx, ly], False, [], [], tol, True, None, None)
    return s, surf


def _uniform(nn, a, b):
    return np.concatenate([np.full(nn, a), np.full(nn, b)])


# ----------------------------------------------------------------------------- Ct and tau
def test_ctau_tilde_and_ce_closed_form_for_uniform_velocity():
    hx, hy = 0.3, 0.2
    s, _ = _grid(3, 2, 3 * hx, 2 * hy)
    E = _q1_closed(hx, hy)
    assert np.allclose(s.Me, E["Me"], rtol=1e-13, atol=1e-16)
    assert np.allclose(s.Le, E["Le"], rtol=1e-13, atol=1e-15)
    a, b = 0.7, -1.3
    U = _uniform(s.nn, a, b)
    Ct = el.Ctau_tilde_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 2)
    Ce = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 2, 1.0)
    for e in range(s.nconn.shape[0]):
        assert np.allclose(Ct[e], _ct_uniform(E, a, b), rtol=1e-13, atol=1e-14)
        assert np.allclose(Ce[e], _ce_uniform(E, a, b), rtol=1e-13, atol=1e-15)
    # Ct(U) is u^T (int grad N_a grad N_b^T) u: symmetric, annihilates constants
    assert np.allclose(Ct[0], Ct[0].T) and np.allclose(Ct[0] @ np.ones(4), 0, atol=1e-14)


def _tau(s, a, b, dt, D, fn=None):
    fn = fn or el.calculate_tau
    U = _uniform(s.nn, a, b)
    Ct = el.Ctau_tilde_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 2)
    Ce = el.conservative_convec_elem_calc(s.nconn, U, s.wq, s.N, s.Nderiv, s.nn, 2, 1.0)
    return fn(s.nconn, U, Ce, Ct, s.N, s.nn, 2, D, dt, r=2)


def _tau_checks(fn=None):
    """True when calculate_tau (or a mutant) passes every hand value."""
    s, _ = _grid(2, 2, 2.0, 2.0)  # unit square elements
    ok = True
    # u = (1, 0), h = 1: |Ce| = sqrt(10)/6, |Ct| = 2 sqrt(10)/6, so tau1 = 1/2; dt = 1 gives
    # tau2 = 1/2; Re = |u|^2 tau1 / (D + 1e-6) = 2, tau3 = 1, tau = (4 + 4 + 1)^(-1/2) = 1/3
    with np.errstate(all="ignore"):
        t = _tau(s, 1.0, 0.0, 1.0, 0.25 - 1e-6, fn)
        ok &= bool(np.allclose(t, 1.0 / 3.0, rtol=1e-13))
        # dt -> 0: tau -> dt / 2
        dt = 1e-9
        t = _tau(s, 0.6, -0.8, dt, 0.01, fn)
        ok &= bool(np.all(np.abs(t / (dt / 2) - 1.0) < 1e-12))
        # D = 0: the 1e-6 floor sets tau3 = h^2 / (4e-6) = 2.5e5, which dominates when
        # |u| = 1e-7 (tau1 = 5e6) and dt = 1e7 (tau2 = 5e6)
        t = _tau(s, 1e-7, 0.0, 1e7, 0.0, fn)
        ok &= bool(np.allclose(t, (2 / 5e6 ** 2 + 1 / 2.5e5 ** 2) ** -0.5, rtol=1e-9))
    return ok


def test_calculate_tau_hand_values():
    assert _tau_checks()


@pytest.mark.parametrize("old,new", [("D = D + 1e-6", "D = D"),
                                     ("** (-1 / r)", "** (-1 / (r + 1))"),
                                     ("t2 = dt / 2", "t2 = dt"),
                                     (".mean(axis=1)", ".sum(axis=1)")])
def test_tau_checks_reject_planted_mistakes(old, new):
    assert not _tau_checks(_mutant(el.calculate_tau, old, new))


# ----------------------------------------------------------------------------- scalar step
def _dense_scalar_step(s, E, un, u, phi_n, dt, D, gd, phi_d, supg):
    """CN step of passive_transport.py:14-61 for uniform old / new velocities, from closed-form
    element matrices, dense assembly and a direct solve."""
    nn = s.nn
    M = _assemble(s.nconn, nn, E["Me"])
    L = _assemble(s.nconn, nn, E["Le"])
    C = _assemble(s.nconn, nn, _ce_uniform(E, *u))
    Cn = _assemble(s.nconn, nn, _ce_uniform(E, *un))
    lhs = M / dt + 0.5 * (C + D * L)
    rhs_m = M / dt - 0.5 * (Cn + D * L)
    if supg:
        # tau from the old velocity (|U_n|^2, Ct(U_n)) and |Ce| of the new-velocity Ce
        ce = np.linalg.norm(_ce_uniform(E, *u))
        Ctn = _ct_uniform(E, *un)
        ct = np.linalg.norm(Ctn)
        t1 = ce / ct
        t3 = t1 * (un[0] ** 2 + un[1] ** 2) * ce / ((D + 1e-6) * ct)
        tau = (t1 ** -2 + (dt / 2) ** -2 + t3 ** -2) ** -0.5
        Mt = _assemble(s.nconn, nn, tau * _ce_uniform(E, *un).T)
        Kt = _assemble(s.nconn, nn, tau * Ctn)
        lhs = lhs + Mt / dt + 0.5 * Kt
        rhs_m = rhs_m + Mt / dt - 0.5 * Kt
    rhs = rhs_m @ phi_n
    lhs[gd, :] = 0.0
    lhs[gd, gd] = 1.0
    rhs[gd] = phi_d
    return np.linalg.solve(lhs, rhs)


def _scalar_step_error(step_fn, supg):
    hx, hy = 0.25, 0.2
    s, surf = _grid(6, 5, 6 * hx, 5 * hy)
    E = _q1_closed(hx, hy)
    un, u = (0.8, -0.3), (0.5, 0.6)  # different directions and magnitudes
    phi_n = RNG.random(s.nn)
    gd = np.array(surf[3])  # left face
    phi_d = np.linspace(1.0, 2.0, len(gd))
    dt, D = 0.05, 0.01
    ref = _dense_scalar_step(s, E, un, u, phi_n, dt, D, gd, phi_d, supg)
    got, _ = step_fn(s, _uniform(s.nn, *un), _uniform(s.nn, *u), phi_n, dt, D, phi_d, gd, supg)
    return ns.rel_l2(got, ref)


@pytest.mark.parametrize("supg", [False, True])
def test_scalar_step_uniform_velocities_dense(supg):
    assert _scalar_step_error(ns.scalar_step, supg) < 1e-11


@pytest.mark.parametrize("old,new", [
    ("calculate_tau(nconn, U_n, Ce,", "calculate_tau(nconn, U, Ce,"),          # tau from U
    ("calculate_tau(nconn, U_n, Ce,", "calculate_tau(nconn, U_n, Ce_n,"),      # |Ce| of the old Ce
    ("Ctau_tilde_elem_calc(nconn, U_n,", "Ctau_tilde_elem_calc(nconn, U,"),    # Ct from U
    ("Ce_n.transpose(0, 2, 1)", "Ce_n"),                                       # M_tau untransposed
    ("D, dt, r=2)", "D, dt, r=1)"),                                            # r exponent
])
def test_scalar_step_check_rejects_planted_supg_mistakes(old, new):
    assert _scalar_step_error(_mutant(ns.scalar_step, old, new), True) > 1e-6


def _conservation_change(step_fn, beta_check=False):
    """Relative change of 1^T M_c phi over one plain CN step on a walled box (no scalar Dirichlet
    nodes) with velocities that vanish on the walls but have nonzero divergence inside."""
    s, surf = _grid(10, 8, 1.0, 0.8)
    wall = np.unique(np.concatenate(surf))
    U_n = RNG.standard_normal(2 * s.nn)
    U = U_n + 0.5 * RNG.standard_normal(2 * s.nn)
    for V in (U_n, U):
        V[wall] = 0.0
        V[s.nn + wall] = 0.0
    phi_n = RNG.random(s.nn)
    phi, _ = step_fn(s, U_n, U, phi_n, 0.02, 0.01, None, None, False)
    M = emult.assemble(s.nconn, s.nn, s.Me)
    m0, m1 = (M @ phi_n).sum(), (M @ phi).sum()
    return abs(m1 - m0) / abs(m0)


def test_plain_step_conserves_mass_with_walls():
    assert _conservation_change(ns.scalar_step) < 1e-12


def test_conservation_check_rejects_beta_half():
    bad = _mutant(ns.scalar_step, "s.N, s.Nderiv, nn, ndim, 1)", "s.N, s.Nderiv, nn, ndim, 0.5)", count=2)
    assert _conservation_change(bad) > 1e-6


def _decay_error(update_fn):
    """U = 0: the nodal mode cos(k pi x / lx) cos(l pi y / ly) decays by
    g = (1/dt - D lam / 2) / (1/dt + D lam / 2), lam the discrete Q1 eigenvalue."""
    nex, ney, lx, ly = 12, 10, 1.2, 0.8
    s, _ = _grid(nex, ney, lx, ly)
    hx, hy = lx / nex, ly / ney
    k, l = 3, 2
    tx, ty = k * np.pi / nex, l * np.pi / ney
    x, y = s.coords[:, 0], s.coords[:, 1]
    phi_n = np.cos(k * np.pi * x / lx) * np.cos(l * np.pi * y / ly)
    lam = 6 / hx ** 2 * (1 - np.cos(tx)) / (2 + np.cos(tx)) + 6 / hy ** 2 * (1 - np.cos(ty)) / (2 + np.cos(ty))
    dt, D = 0.01, 0.05
    g = (1 / dt - 0.5 * D * lam) / (1 / dt + 0.5 * D * lam)
    Z = np.zeros((s.nconn.shape[0], 4, 4))
    phi, _ = update_fn(phi_n, s.nconn, s.Me, Z, Z, s.Le, D, dt, s.nn, None, None, 1e-13)
    return ns.rel_l2(phi, g * phi_n)


def test_plain_step_diffusion_mode_decay():
    assert _decay_error(ns.update_phi) < 1e-11


@pytest.mark.parametrize("old,new", [
    ("- 0.5 * emult.scalar_visc_conv_mult(phi_n", "+ 0.5 * emult.scalar_visc_conv_mult(phi_n"),
    ("+ 0.5 * emult.scalar_visc_conv_mult(phi, Ce_all", "+ emult.scalar_visc_conv_mult(phi, Ce_all"),
    ("res = ((1 / dt) * emult", "res = (dt * emult"),
])
def test_decay_check_rejects_planted_mistakes(old, new):
    assert _decay_error(_mutant(ns.update_phi, old, new)) > 1e-6


# ----------------------------------------------------------------------------- Neumann V-cycle
def _interp_1d(nc):
    """Linear interpolation from nc coarse nodes to 2 nc - 1 fine nodes."""
    I = np.zeros((2 * nc - 1, nc))
    for i in range(nc):
        I[2 * i, i] = 1.0
    for i in range(nc - 1):
        I[2 * i + 1, i] = I[2 * i + 1, i + 1] = 0.5
    return I


def _dense_neumann_vcycle(r, grids):
    """grids: fine to coarse list of (nex, ney, hx, hy).  Pure-Neumann V-cycle of
    multigrid.py:231-330 in dense form: P is the mean projector of each level."""
    Ls, Is = [], []
    for (nex, ney, hx, hy) in grids:
        _, nconn, *_ = ms.create_mesh_2d(nex, ney, nex * hx, ney * hy)
        Ls.append(_assemble(nconn, (nex + 1) * (ney + 1), _q1_closed(hx, hy)["Le"]))
    for (nex, ney, _, _), (nexc, neyc, _, _) in zip(grids[:-1], grids[1:]):
        Is.append(np.kron(_interp_1d(neyc + 1), _interp_1d(nexc + 1)))

    def P(v):
        return v - v.mean()

    def jacobi(x, rhs, L):
        d = np.diag(L)
        return P(2 / 3 * (rhs - (L - np.diag(d)) @ x) / d + 1 / 3 * x)

    N = len(grids)
    rs, xs = [r], [None] * N
    for n in range(N - 1):
        rh = P(rs[n])
        xh = P(2 / 3 * rh / np.diag(Ls[n]))
        rH = P(Is[n].T @ P(rh - Ls[n] @ xh))
        xs[n] = xh
        rs.append(rH)
    xs[N - 1] = jacobi(np.zeros_like(rs[N - 1]), P(rs[N - 1]), Ls[N - 1])
    for n in range(N - 2, -1, -1):
        xs[n] = jacobi(xs[n] + P(Is[n] @ xs[n + 1]), rs[n], Ls[n])
    return xs[0]


def _vcycle_error(vcycle_fn):
    s, *_ = ns.build([8, 8], [1.6, 1.2], False, [], [], 1e-12, False, "multigrid", [2, 2])
    grids = [(8, 8, 0.2, 0.15), (4, 4, 0.4, 0.3), (2, 2, 0.8, 0.6)]
    r = RNG.standard_normal(s.nn) + 0.7  # a nonzero mean, so every projection matters
    return ns.rel_l2(vcycle_fn(r, s.levels, 1e-12, False, False), _dense_neumann_vcycle(r, grids))


def test_neumann_vcycle_dense_three_levels():
    assert _vcycle_error(mg.v_cycle) < 1e-13


@pytest.mark.parametrize("old,new", [
    ("rh = r_lv[n] - r_lv[n].mean()", "rh = r_lv[n]"),
    ("x_lv[n] = single_jacobi_iter(xh, r_lv[n],", "x_lv[n] = single_jacobi_iter(xh, r_lv[n] - r_lv[n].mean(),"),
    ("rN = r_lv[N - 1] - r_lv[N - 1].mean()", "rN = r_lv[N - 1] + 1.0"),
])
def test_neumann_vcycle_check_rejects_planted_mistakes(old, new):
    """Dropping the fine-level projection of r, projecting the stored residual in the up-leg
    sweep, or shifting the coarse residual changes the output."""
    assert _vcycle_error(_mutant(mg.v_cycle, old, new)) > 1e-8


@pytest.mark.parametrize("old,new", [
    ("xh = xh - xh.mean()", "xh = xh"),
    ("rH = rH - rH.mean()\n            rH = _restrict", "rH = rH\n            rH = _restrict"),
    ("rH = _restrict(rH, L, Lc, periodic)\n            rH = rH - rH.mean()",
     "rH = _restrict(rH, L, Lc, periodic)\n            rH = rH"),
    ("xH = xH - xH.mean()", "xH = xH"),
])
def test_neumann_vcycle_projections_that_cannot_change_the_output(old, new):
    """Four of the reference's projections act on vectors whose mean is already zero, or add a
    constant that a later projection removes, so the output is the same to round-off without
    them: 1^T (r - L x) = 1^T r because L 1 = 0 and L is symmetric; 1^T R v = (I 1)^T v = 1^T v
    because interpolation keeps constants; a constant c added to x before a sweep comes out as
    c (Le_off c = -L_diag c) and the sweep's final projection removes it."""
    assert _vcycle_error(_mutant(mg.v_cycle, old, new)) < 1e-13
