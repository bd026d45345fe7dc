"""GPU parity at the benchmark's own size: the 1024x1024 Q1 cavity (1,050,625 nodes) and the
1024x1024 pure-Neumann Poisson solve, against the CPU oracle on the same seeded inputs.

At this size the V-cycle runs its launch shapes of the large levels only: the fused down leg with
4 coarse rows per block on the two fine levels, the LDS-staged up-leg Jacobi on level 1 (at least
2^17 nodes) and the paired small-level kernels from 257^2 down.

Tolerances:
* FP64 form: fields 1e-10 relative L2 with identical iteration counts (as at the smaller sizes).
* Mixed form (FP32 V-cycle inside FP64 flexible CG, FP32 stored convection matrices, FP64
  Krylov vectors, dots and projection): each field (u, v, p) within MIXED_FIELD_TOL = 1e-8
  relative L2 of the FP64 oracle after each step, i.e. within the 1e-8 solver tolerance;
  pressure solves converge to the same 1e-8 relative residual.
"""
import numpy as np
import pytest

from oracle import emult as oem
from oracle import ns

pytestmark = pytest.mark.gpu

N = 1024
TOL = 1e-8
RHO, MU, DT = 1.0, 1e-3, 5e-4
STEPS = 2
MIXED_FIELD_TOL = 1e-8


def _np(x):
    return x.detach().cpu().numpy()


def _fields(U, P, nn):
    return {"u": U[:nn], "v": U[nn:], "p": P}


@pytest.fixture(scope="module")
def oracle_cavity():
    """The oracle's first STEPS steps of the 1024^2 cavity (the benchmark's workload):
    lid u = 1 (corners excluded), Re = 1000, dt = 5e-4, U0 = N(0, 1e-3) from seed 1 with the
    Dirichlet values applied, P0 = 0, MG 1024 -> 16 (7 levels), tol 1e-8."""
    print(f"\n[case] oracle 1024^2 cavity, {STEPS} steps", flush=True)
    s, surf, _, oUd = ns.build([N, N], [1.0, 1.0], False, [0, 1, 2, 3], [], TOL, True, "multigrid", [16, 16])
    oUd = ns.cavity_lid(s, surf, oUd)
    U0 = np.random.default_rng(1).normal(0, 1e-3, 2 * s.nn)
    U0[s.gamma_d] = oUd
    P0 = np.zeros(s.nn)
    U, P = U0.copy(), P0.copy()
    out = []
    for _ in range(STEPS):
        U, P, ui, pi = ns.four_step_timestep(s, U, P, DT, RHO, MU, oUd, 0, 0)
        out.append((U.copy(), P.copy(), (ui, pi)))
    return U0, P0, s.nn, out


def _gpu_cavity(precision, ce_format):
    import jax_fem_cfd.setup.simulation as sim
    cfg = sim.Config(mesh="simple", solver="standard",
                     mesh_config=sim.SimpleMeshConfig([N, N], [1.0, 1.0], 1, False),
                     solver_config=sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, TOL, True, "multigrid",
                                                            [16, 16], precision=precision, ce_format=ce_format))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    U_d = sim.set_face_velocity(U_d, surf[2][1:-1], gamma_d, 1.0, nn, 1)
    return timestep, U_d


def test_cavity_1024_fp64_steps_match_oracle(oracle_cavity):
    U0, P0, nn, ref = oracle_cavity
    print("[case] GPU 1024^2 cavity FP64", flush=True)
    timestep, U_d = _gpu_cavity("fp64", None)
    U, P = U0, P0
    for k, (Uo, Po, its) in enumerate(ref):
        U, P, ui, pi = timestep(U, P, DT, RHO, MU, U_d, 0, 0)
        Ug, Pg = _np(U), _np(P)
        err = {f: ns.rel_l2(a, b) for (f, a), b in zip(_fields(Ug, Pg, nn).items(), _fields(Uo, Po, nn).values())}
        print(f"  step {k + 1}: iters GPU ({ui},{pi}) oracle {its}  " +
              "  ".join(f"{f} {e:.2e}" for f, e in err.items()), flush=True)
        assert (ui, pi) == its
        assert max(err.values()) < 1e-10


@pytest.mark.parametrize("ce_format", ["f32"])
def test_cavity_1024_mixed_steps_within_stated_tolerance(oracle_cavity, ce_format):
    U0, P0, nn, ref = oracle_cavity
    print(f"[case] GPU 1024^2 cavity mixed (Ce {ce_format})", flush=True)
    timestep, U_d = _gpu_cavity("mixed", ce_format)
    U, P = U0, P0
    for k, (Uo, Po, its) in enumerate(ref):
        U, P, ui, pi = timestep(U, P, DT, RHO, MU, U_d, 0, 0)
        Ug, Pg = _np(U), _np(P)
        err = {f: ns.rel_l2(a, b) for (f, a), b in zip(_fields(Ug, Pg, nn).items(), _fields(Uo, Po, nn).values())}
        print(f"  step {k + 1}: iters GPU ({ui},{pi}) oracle {its}  " +
              "  ".join(f"{f} {e:.2e}" for f, e in err.items()), flush=True)
        assert max(err.values()) < MIXED_FIELD_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_neumann_poisson_1024(precision):
    """The Poisson solve alone at 1024^2: pure Neumann, rhs N(0,1) from seed 0 projected to zero
    mean, x0 = 0, MG 1024 -> 16, PCG to 1e-8.  FP64: the oracle's iteration count and its solution
    to 1e-10.  Mixed (FP32 V-cycle, FP64 flexible CG): converged, the true FP64 residual
    |b - L x| / |b| <= 1e-8 computed here on the CPU, and the zero-mean solution within
    MIXED_FIELD_TOL of the oracle's."""
    from jax_fem_cfd._context import Context
    from oracle.solvers import cg
    print(f"\n[case] Neumann Poisson 1024^2 {precision}", flush=True)
    s, *_ = ns.build([N, N], [1.0, 1.0], False, [], [], TOL, False, "multigrid", [16, 16])
    b = np.random.default_rng(0).standard_normal(s.nn)
    b -= b.mean()

    def A(x):
        return oem.laplacian_element_mult(x, s.Le, s.nconn, s.nn)

    xr, kr = cg(A, b, tol=TOL, maxiter=1000, M=s.M_op)
    c = Context(N, N, 1.0, 1.0, False, tol=TOL, set_zeroP=False, precond="multigrid", mg_levels=len(s.levels),
                precision=precision)
    x, k, conv = c.pressure_cg(b, tol=TOL, maxiter=1000)
    xg = _np(x)
    res = np.linalg.norm(b - A(xg)) / np.linalg.norm(b)
    err = ns.rel_l2(xg - xg.mean(), xr - xr.mean())
    print(f"  iterations {k} oracle {kr}  true residual {res:.2e}  rel x {err:.2e}", flush=True)
    assert conv and res <= TOL
    if precision == "fp64":
        assert k == kr and ns.rel_l2(xg, xr) < 1e-10
    else:
        assert err < MIXED_FIELD_TOL and ns.rel_l2(xg, xr) < MIXED_FIELD_TOL


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
def test_vcycle_1024_pinned(precision, tol):
    """One pinned V-cycle on the 1024^2 hierarchy (7 levels to 17^2) against the oracle's: the
    launch shapes of the benchmark's preconditioner, including the up leg that runs levels 3, 2
    and 1 in one kernel.  FP64: 1e-12 relative L2 (the same bound as the small-grid V-cycle test);
    mixed (FP32 levels, exact FP32 coarse inverse against the oracle's coarse CG): 2e-5."""
    from jax_fem_cfd._context import Context
    s, *_ = ns.build([N, N], [1.0, 1.0], False, [], [], 1e-12, True, "multigrid", [16, 16])
    print(f"\n[case] 1024^2 pinned V-cycle, {precision}", flush=True)
    c = Context(N, N, 1.0, 1.0, False, tol=1e-12, set_zeroP=True, precond="multigrid", mg_levels=len(s.levels),
                precision=precision, gamma_p=[0], mg_gamma_p=[L["gamma_p"] for L in s.levels])
    r = np.random.default_rng(7).standard_normal(s.nn)
    r[0] = 0
    ref = s.M_op(r)
    got = _np(c.vcycle(r))
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < tol, err
