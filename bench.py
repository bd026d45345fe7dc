"""Benchmark: full 4-step timesteps/s on the 1024x1024 Q1 lid-driven cavity.

Workload (one "step" = one four_step_timestep): 2D walled unit box, 1024x1024
Q1 elements, lid u = 1 (corners excluded), no-slip elsewhere, Re = 1000
(rho = 1, mu = 1e-3), U0 = N(0, 1e-3) noise from seed 1 with the Dirichlet
values applied, P0 = 0, dt = 5e-4; BiCGSTAB momentum + geometric-multigrid
PCG pressure (1024^2 -> 16^2, pinned pressure), both to 1e-8.

`value` is the mixed-precision form (FP32 multigrid inside an FP64 flexible
CG, FP32 stored convection matrices, FP64 Krylov vectors, dots and
projection).  The FP64 form (the reference's algorithm) runs on the same GPU
in the same job and is reported beside it, with the per-field difference of
the two after the timed steps.  Also reported: the pure-Neumann 1024^2
Poisson solve alone (solves/s, FP64 and mixed, 1e-8), the roofline of the
dominant kernel (momentum operator apply) and the CPU oracle timed on this
host on a bounded sample.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "jax-fem-cfd_amd"))

import numpy as np  # noqa: E402

N_EL = 1024
RHO, MU, DT, TOL = 1.0, 1e-3, 5e-4, 1e-8
HBM_PEAK_GBS = 8000.0


def _cavity(precision, ce_format, n=N_EL):
    import jax_fem_cfd.setup.simulation as sim
    cfg = sim.Config(mesh="simple", solver="standard",
                     mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                     solver_config=sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, TOL, True, "multigrid",
                                                            [16, 16], precision=precision, ce_format=ce_format))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
    U_d = sim.set_face_velocity(U_d, surf[2][1:-1], gamma_d, 1.0, nn, 1)
    U0 = np.random.default_rng(1).normal(0, 1e-3, 2 * nn)
    U0[gamma_d] = U_d
    return compute, timestep, U_d, U0, np.zeros(nn), nn


def run_form(precision, ce_format, steps, warmup):
    """`warmup` untimed steps through the native step loop (its buffers and graphs are then in
    place), then `steps` timed steps through the same loop.  The same `steps` steps are then run
    again from the same start state with one Python `timestep` call each (the reference's
    timestep_loop shape) and timed beside it."""
    import torch
    compute, timestep, U_d, U0, P0, nn = _cavity(precision, ce_format)
    F_p, F = compute(U_d, None, 0, 0)
    U = torch.as_tensor(U0, device="cuda")
    P = torch.as_tensor(P0, device="cuda")
    Ud = torch.as_tensor(U_d, device="cuda")
    ctx = timestep.context
    if warmup > 0:
        U, P, _ = ctx.four_step_rollout(U, P, warmup, DT, RHO, MU, Ud, F_p, F)
    Uw, Pw = U.clone(), P.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U, P, its = ctx.four_step_rollout(U, P, steps, DT, RHO, MU, Ud, F_p, F)
    torch.cuda.synchronize()
    dt_s = time.perf_counter() - t0
    U2, P2 = Uw, Pw
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        U2, P2, ui, pi = timestep(U2, P2, DT, RHO, MU, Ud, F_p, F)
    torch.cuda.synchronize()
    py_s = (time.perf_counter() - t1) / steps
    return dict(sec=dt_s, its=[tuple(map(int, r)) for r in its], U=U, P=P, Uw=Uw, Pw=Pw, ctx=ctx, nn=nn,
                py_step_s=py_s)


def poisson_alone(precision, solves=3):
    """Pure-Neumann 1024^2 Poisson: rhs N(0,1) seed 0 projected to zero mean, x0 = 0,
    MG 1024^2 -> 16^2, PCG to 1e-8 relative residual."""
    import torch
    from jax_fem_cfd._context import Context
    n = N_EL
    c = Context(n, n, 1.0, 1.0, False, tol=TOL, set_zeroP=False, precond="multigrid", mg_levels=7,
                precision=precision)
    b = np.random.default_rng(0).standard_normal(c.nn)
    b -= b.mean()
    bt = torch.as_tensor(b, device="cuda")
    x, k, conv = c.pressure_cg(bt, tol=TOL)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(solves):
        x, k, conv = c.pressure_cg(bt, tol=TOL)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / solves
    # true FP64 residual through the native Laplacian
    Le = c.element_constant("Le")
    Ax = c.elem_mult_const(Le, x)
    res = float(torch.linalg.norm(bt - Ax) / torch.linalg.norm(bt))
    return dict(solves_per_s=1.0 / sec, ms_per_solve=sec * 1e3, iters=k, converged=conv, true_rel_residual=res)


def fourier_velocity(coords, seed, kmax=8):
    """Divergence-free velocity whose modes |k| <= kmax have amplitudes N(0, 1) / |k|^2 (seeded):
    stream function psi = sum_{m, n >= 1, m^2 + n^2 <= kmax^2} a_mn sin(m pi x) sin(n pi y) / (pi |k|^3),
    a_mn ~ N(0, 1), |k| = sqrt(m^2 + n^2); U = (d psi / dy, -d psi / dx), so each mode's velocity
    amplitude is a_mn (n or m) / |k|^3 <= |a_mn| / |k|^2."""
    x, y = coords[:, 0], coords[:, 1]
    rng = np.random.default_rng(seed)
    u = np.zeros_like(x)
    v = np.zeros_like(x)
    for m in range(1, kmax + 1):
        for n in range(1, kmax + 1):
            k2 = m * m + n * n
            if k2 > kmax * kmax:
                continue
            a = rng.standard_normal() / (np.pi * k2 ** 1.5)
            u += a * n * np.pi * np.sin(m * np.pi * x) * np.cos(n * np.pi * y)
            v -= a * m * np.pi * np.cos(m * np.pi * x) * np.sin(n * np.pi * y)
    return np.concatenate([u, v])


def gradient_256(steps=200, checkpoint=20):
    """Config 5: 256^2 walled unit box, rho = 1, mu = 1e-3, kappa = 1e-4, dt = 1e-3, plain scalar
    step, MG 256^2 -> 16^2, solves to 1e-10.  phi0 = 1 in a seeded rectangle (seed 2), 0 elsewhere;
    the target is the solver's phi after `steps` steps from fourier_velocity(seed 5); the
    gradient is taken at U0 = fourier_velocity(seed 6) (walls set to 0) of
    L = 1/2 |phi_steps - phi_target|^2 (max |U0| ~ 0.5: CFL ~ 0.1 at dt = 1e-3, h = 1/256), by the
    hand-written adjoint with checkpoints every
    `checkpoint` steps.  Reported: wall time of one value-and-gradient, and the central finite
    difference of L along one random direction (seed 7) against <grad, d>."""
    import torch
    import jax_fem_cfd.setup.simulation as sim
    n, dt, rho, mu, D = 256, 1e-3, 1.0, 1e-3, 1e-4
    cfg = sim.Config(mesh="simple", solver="passive_transport",
                     mesh_config=sim.SimpleMeshConfig([n, n], [1.0, 1.0], 1, False),
                     solver_config=sim.PassiveTransportConfig(
                         sim.StandardSolverConfig([0, 1, 2, 3], [], [], False, 1e-10, True, "multigrid", [16, 16]),
                         sim.ADEConfig([], False)))
    compute, timestep, coords, surf, h, nn, nnx, nny, gamma_d, U_d, gdp = sim.setup(cfg)
    rng = np.random.default_rng(2)
    x0, y0 = rng.uniform(0.1, 0.5, 2)
    w, hgt = rng.uniform(0.2, 0.4, 2)
    x, y = coords[:, 0], coords[:, 1]
    phi0 = ((x >= x0) & (x <= x0 + w) & (y >= y0) & (y <= y0 + hgt)).astype(float)
    dev = torch.device("cuda", 0)
    P0 = torch.zeros(nn, dtype=torch.float64, device=dev)
    phi0t = torch.tensor(phi0, dtype=torch.float64, device=dev)

    def vel(seed):
        U = fourier_velocity(coords, seed)
        U[gamma_d] = U_d
        return U

    def run(U):
        return sim.timestep_loop_scan(dt, steps, compute, timestep, rho, mu, D, U_d, None, U, P0, phi0t, None,
                                      checkpoint=checkpoint)[2]

    with torch.no_grad():
        target = run(torch.tensor(vel(5), dtype=torch.float64, device=dev))
    U0 = torch.tensor(vel(6), dtype=torch.float64, device=dev, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = 0.5 * torch.sum((run(U0) - target) ** 2)
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = U0.grad.detach()
    d = torch.tensor(np.random.default_rng(7).standard_normal(2 * nn), dtype=torch.float64, device=dev)
    eps = 1e-4
    with torch.no_grad():
        lp = 0.5 * torch.sum((run(U0.detach() + eps * d) - target) ** 2).item()
        lm = 0.5 * torch.sum((run(U0.detach() - eps * d) - target) ** 2).item()
    fd = (lp - lm) / (2 * eps)
    an = float(torch.dot(g, d))
    return {"grid": "256x256 Q1", "steps": steps, "checkpoint": checkpoint, "loss": loss.item(),
            "seconds_value_and_grad": t2 - t0, "seconds_forward": t1 - t0, "seconds_backward": t2 - t1,
            "fd_directional": fd, "adjoint_directional": an, "fd_rel_err": abs(fd - an) / max(abs(an), 1e-300),
            "params": 2 * nn}


def cpu_baseline(U_start, P_start, first_step, steps=3, threads=None):
    """The oracle (NumPy FP64 restatement) on this host on a bounded sample of the timed window:
    it starts from the FP64 GPU state after step `first_step - 1` (the oracle and the FP64 GPU
    form agree to 1e-10 there), runs step `first_step` untimed (warm-up), then times the next
    `steps` steps.  OMP / OpenBLAS / MKL threads = `threads` (default: the cores this process
    may run on); NumPy's gathers and scatters run on one thread whatever that setting."""
    import tempfile
    cores = len(os.sched_getaffinity(0))
    threads = cores if threads is None else threads
    fd, path = tempfile.mkstemp(suffix=".npz")
    os.close(fd)
    np.savez(path, U=U_start, P=P_start)
    code = f"""
import sys, time, json
sys.path.insert(0, {ROOT!r})
import numpy as np
from oracle import ns
s, surf, h, U_d = ns.build([{N_EL}, {N_EL}], [1.0, 1.0], False, [0, 1, 2, 3], [], {TOL}, True, "multigrid", [16, 16])
U_d = ns.cavity_lid(s, surf, U_d)
z = np.load({path!r})
U, P = z["U"], z["P"]
U, P, ui, pi = ns.four_step_timestep(s, U, P, {DT}, {RHO}, {MU}, U_d, 0, 0)
t0 = time.perf_counter()
its = []
for _ in range({steps}):
    U, P, ui, pi = ns.four_step_timestep(s, U, P, {DT}, {RHO}, {MU}, U_d, 0, 0)
    its.append((ui, pi))
sec = time.perf_counter() - t0
print(json.dumps(dict(sec=sec, steps={steps}, its=its)))
"""
    t = str(threads)
    env = dict(os.environ, OMP_NUM_THREADS=t, OPENBLAS_NUM_THREADS=t, MKL_NUM_THREADS=t)
    try:
        out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=900)
    finally:
        os.unlink(path)
    if out.returncode != 0:
        return dict(error=out.stderr[-400:])
    r = json.loads(out.stdout.strip().splitlines()[-1])
    r.update(value=r["steps"] / r["sec"], cores=cores, threads=threads,
             window=[first_step + 1, first_step + steps])
    return r


def taylor_green_64(with_cpu):
    """Config 1: the 2D periodic Taylor-Green vortex on [0, 2 pi]^2, 64x64 Q1 (65^2 nodes, periodic),
    nu = 0.01, dt = 1e-3, 100 steps of the 4-step projection from t = 0, FP64, MG-PCG and BiCGSTAB
    tol 1e-10 (examples/taylor_green.py).  GPU: 100 steps through the native step loop after an
    untimed 10-step warm-up from the same state, with the error of the fields against the exact
    decaying vortex.  CPU (the cpu_baseline leg): the oracle runs the same 100 steps from the same
    state in a subprocess, and the GPU fields are compared with its fields."""
    import tempfile
    import torch
    import jax_fem_cfd.setup.simulation as sim
    n, L, rho, mu, dt, steps = 64, 2 * np.pi, 1.0, 0.01, 1e-3, 100
    cfg = sim.Config(mesh="simple", solver="standard",
                     mesh_config=sim.SimpleMeshConfig([n, n], [L, L], 1, True),
                     solver_config=sim.StandardSolverConfig([], [], [], False, 1e-10, False, "multigrid", [2, 2]))
    compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)

    def exact(t):
        x, y = coords[:, 0], coords[:, 1]
        f = np.exp((-2 * mu / rho) * t)
        u, v = np.sin(x) * np.cos(y) * f, -np.cos(x) * np.sin(y) * f
        p = (rho / 4) * (np.cos(2 * x) + np.cos(2 * y)) * f * f
        return np.concatenate([u, v]), p - p.mean()

    U0, P0 = exact(0.0)
    F_p, F = compute(U_d, None, 0, 0)
    ctx = timestep.context
    U = torch.as_tensor(U0, device="cuda")
    P = torch.as_tensor(P0, device="cuda")
    ctx.four_step_rollout(U, P, 10, dt, rho, mu, None, F_p, F)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Ug, Pg, its = ctx.four_step_rollout(U, P, steps, dt, rho, mu, None, F_p, F)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    Ug, Pg = Ug.cpu().numpy(), Pg.cpu().numpy()
    Ue, Pe = exact(steps * dt)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    out = {"workload": "2D periodic Taylor-Green 64x64 Q1, nu=0.01, dt=1e-3, 100 steps, FP64, tol 1e-10",
           "timesteps_per_s": steps / sec, "ms_per_step": 1e3 * sec / steps,
           "mean_iters": [float(np.mean([i[0] for i in its])), float(np.mean([i[1] for i in its]))],
           "rel_l2_vs_exact": {"u": rel(Ug[:nn], Ue[:nn]), "v": rel(Ug[nn:], Ue[nn:]),
                               "p": rel(Pg - Pg.mean(), Pe)}}
    if not with_cpu:
        return out
    fd, path = tempfile.mkstemp(suffix=".npz")
    os.close(fd)
    np.savez(path, U=U0, P=P0)
    code = f"""
import sys, time, json
sys.path.insert(0, {ROOT!r})
import numpy as np
from oracle import ns
s, *_ = ns.build([{n}, {n}], [{L!r}, {L!r}], True, [], [], 1e-10, False, "multigrid", [2, 2])
z = np.load({path!r})
U, P = z["U"], z["P"]
t0 = time.perf_counter()
its = []
for _ in range({steps}):
    U, P, ui, pi = ns.four_step_timestep(s, U, P, {dt}, {rho}, {mu}, None, 0, 0)
    its.append((ui, pi))
sec = time.perf_counter() - t0
np.savez({path!r}, U=U, P=P)
print(json.dumps(dict(sec=sec, steps={steps}, its=its)))
"""
    cores = len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OPENBLAS_NUM_THREADS=str(cores), MKL_NUM_THREADS=str(cores))
    try:
        res = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
        if res.returncode != 0:
            out["cpu_baseline"] = dict(error=res.stderr[-400:])
            return out
        r = json.loads(res.stdout.strip().splitlines()[-1])
        z = np.load(path)
        Uo, Po = z["U"], z["P"]
    finally:
        os.unlink(path)
    out["cpu_baseline"] = {"kind": "port", "timesteps_per_s": r["steps"] / r["sec"], "cores": cores,
                           "iters_equal": [tuple(map(int, a)) for a in its] == [tuple(a) for a in r["its"]],
                           "gpu_vs_oracle_rel_l2": {"u": rel(Ug[:nn], Uo[:nn]), "v": rel(Ug[nn:], Uo[nn:]),
                                                    "p": rel(Pg, Po)}}
    return out


def roofline_only(reps):
    """The dominant kernel alone (for counter runs): one mixed-form timestep to build the FP32
    Ce, `reps` momentum operator applies, then a plain device copy of known size (the
    calibration of the HBM byte counters)."""
    import torch
    mix = run_form("mixed", "f32", 1, 0)
    ms = mix["ctx"].bench_momentum_apply(reps)
    n_copy = 1 << 25  # 32 Mi doubles = 256 MiB read + 256 MiB written
    x = torch.ones(n_copy, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    ne, nn = N_EL * N_EL, mix["nn"]
    print(json.dumps({"kernel_us": ms * 1e3, "reps": reps, "bytes_per_launch": 16 * ne * 4 + 6 * nn * 8,
                      "copy_bytes_each_way": n_copy * 8}))


def channel_3d(steps=200, warmup=3):
    """Config 4 (Q1 hex8): 128x64x64 box [0,2]x[0,1]x[0,1], parabolic inflow
    u = 16 y(1-y) z(1-z) (u_max = 1) on the left face, outlet (p = 0) on the right face, no-slip on
    the other four, rho = 1, mu = 1e-3, dt = 2e-3, BiCGSTAB + PCG, tol 1e-8, from rest.  Pressure
    preconditioners: the approximate inverse (the reference's 3D choice: it has no 3D multigrid)
    and the box-grid V-cycle (128x64x64 -> 8x4x4, 5 levels, outlet pins on every level).  FP64
    form beside the mixed form (FP32-stored Ce); per-field relative L2 of mixed vs FP64 after
    `steps` steps."""
    import torch
    import jax_fem_cfd.setup.simulation as sim
    n, L = [128, 64, 64], [2.0, 1.0, 1.0]

    def run(prec, precond, final):
        cfg = sim.Config(mesh="simple", solver="standard", mesh_config=sim.SimpleMeshConfig(n, L, 1, False),
                         solver_config=sim.StandardSolverConfig([0, 2, 3, 4, 5], [], [1], False, 1e-8, True,
                                                                precond, final, prec))
        compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
        nodes = surf[3]
        y, z = coords[nodes, 1], coords[nodes, 2]
        U_d = sim.set_face_velocity(U_d, nodes, gamma_d, 16 * y * (1 - y) * z * (1 - z), nn, 1)
        U0 = np.zeros(3 * nn)
        U0[np.asarray(gamma_d)] = U_d
        U = torch.tensor(U0, dtype=torch.float64, device="cuda")
        P = torch.zeros(nn, dtype=torch.float64, device="cuda")
        F_p, F = compute(U_d, None, 0, 0)
        for _ in range(warmup):
            timestep(U, P, 2e-3, 1.0, 1e-3, U_d, F_p, F)
        torch.cuda.synchronize()
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            U, P, ui, pi = timestep(U, P, 2e-3, 1.0, 1e-3, U_d, F_p, F)
            its.append((ui, pi))
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        res = {"timesteps_per_s": steps / sec, "ms_per_step": 1e3 * sec / steps,
               "mean_iters": [float(np.mean([i[0] for i in its])), float(np.mean([i[1] for i in its]))]}
        return res, (U.cpu().numpy(), P.cpu().numpy(), nn)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    def compare(f64, fm):
        (U64, P64, nn), (Um, Pm, _) = f64, fm
        d = {c: rel(Um[i * nn:(i + 1) * nn], U64[i * nn:(i + 1) * nn]) for i, c in enumerate("uvw")}
        d["p"] = rel(Pm, P64)
        return d

    out = {}
    fields = {}
    for prec in ("fp64", "mixed"):
        out[prec], fields[prec] = run(prec, "approx_inv", None)
    nn = fields["fp64"][2]
    out["mixed_vs_fp64_rel_l2"] = compare(fields["fp64"], fields["mixed"])
    out["speedup_mixed_vs_fp64"] = out["mixed"]["timesteps_per_s"] / out["fp64"]["timesteps_per_s"]
    mgf = {}
    mg = {"final_multigrid_mesh": [8, 4, 4], "mg_levels": 5}
    for prec in ("fp64", "mixed"):
        mg[prec], mgf[prec] = run(prec, "multigrid", [8, 4, 4])
    mg["mixed_vs_fp64_rel_l2"] = compare(mgf["fp64"], mgf["mixed"])
    # the two preconditioners solve the same systems to the same tolerance
    mg["multigrid_vs_approx_inv_fp64_rel_l2"] = compare(fields["fp64"], mgf["fp64"])
    out["multigrid"] = mg
    out["config"] = {"workload": "3D channel 128x64x64 Q1 hex8, approx_inv PCG (and box-grid multigrid PCG), "
                                 f"dt=2e-3, {steps} steps from rest",
                     "nodes": int(nn), "elements": 128 * 64 * 64, "steps": steps}
    return out


def channel_3d_q2(warmup=1):
    """Config 4 on triquadratic (hex27) elements: the 128x64x64 channel of channel_3d with Q2
    velocity and pressure (257x129x129 = 4,276,737 nodes, 27x27 Ce = 3.06 GB FP64 / 1.53 GB FP32),
    tol 1e-8, from rest, FP64 and mixed forms.  Pressure preconditioners: approx_inv (3 timed
    steps, about 0.7 s each) and the box-grid V-cycle on hex27 levels (128x64x64 -> 8x4x4, 5 timed
    steps); `warmup` untimed steps first."""
    import torch
    import jax_fem_cfd.setup.simulation as sim
    n, L = [128, 64, 64], [2.0, 1.0, 1.0]

    def run(prec, precond, final, steps):
        cfg = sim.Config(mesh="simple", solver="standard", mesh_config=sim.SimpleMeshConfig(n, L, 2, False),
                         solver_config=sim.StandardSolverConfig([0, 2, 3, 4, 5], [], [1], False, 1e-8, True,
                                                                precond, final, prec))
        compute, timestep, coords, surf, h, nn, gamma_d, U_d = sim.setup(cfg)
        nodes = surf[3]
        y, z = coords[nodes, 1], coords[nodes, 2]
        U_d = sim.set_face_velocity(U_d, nodes, gamma_d, 16 * y * (1 - y) * z * (1 - z), nn, 1)
        U0 = np.zeros(3 * nn)
        U0[np.asarray(gamma_d)] = U_d
        U = torch.tensor(U0, dtype=torch.float64, device="cuda")
        P = torch.zeros(nn, dtype=torch.float64, device="cuda")
        F_p, F = compute(U_d, None, 0, 0)
        for _ in range(warmup):
            U, P, _, _ = timestep(U, P, 2e-3, 1.0, 1e-3, U_d, F_p, F)
        torch.cuda.synchronize()
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            U, P, ui, pi = timestep(U, P, 2e-3, 1.0, 1e-3, U_d, F_p, F)
            its.append((ui, pi))
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        return {"timesteps_per_s": steps / sec, "ms_per_step": 1e3 * sec / steps, "steps": steps,
                "mean_iters": [float(np.mean([i[0] for i in its])), float(np.mean([i[1] for i in its]))]}

    out = {}
    for precond, final, steps in (("approx_inv", None, 3), ("multigrid", [8, 4, 4], 5)):
        out[precond] = {prec: run(prec, precond, final, steps) for prec in ("fp64", "mixed")}
    out["config"] = {"workload": "3D channel 128x64x64 hex27 (Q2), approx_inv and multigrid PCG, dt=2e-3, "
                                 f"timed from step {warmup + 1} after rest",
                     "nodes": 257 * 129 * 129, "elements": 128 * 64 * 64}
    return out


def copy_bandwidth(n_bytes=1 << 30, reps=10):
    """Device-to-device copy of n_bytes (read + write counted), timed with events on torch's
    stream: the practical streaming rate the roofline kernel is also compared against."""
    import torch
    x = torch.ones(n_bytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    del x, y
    return 2 * n_bytes / sec / 1e9


def _traffic_from_profile():
    """HBM bytes per launch of the roofline kernel from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "momentum_apply_1024_counters.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU-oracle steps (after one untimed)")
    ap.add_argument("--no-grad", action="store_true", help="skip the 256^2 adjoint-gradient measurement")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D channel measurement")
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the dominant kernel and a calibration copy (for counter runs)")
    ap.add_argument("--only-3d", action="store_true",
                    help="run only the 3D channel measurements (hex8 and hex27) and print them")
    ap.add_argument("--steps-3d", type=int, default=200, help="timed steps of each hex8 channel run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the timestep is a single-GPU path; "
              "multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)

    if args.roofline_only:
        roofline_only(20)
        return
    if args.only_3d:
        print(json.dumps({"channel_3d": channel_3d(args.steps_3d), "channel_3d_q2": channel_3d_q2()}))
        return
    f64 = run_form("fp64", "f64", args.steps, args.warmup)
    mix = run_form("mixed", "f32", args.steps, args.warmup)
    cpu_start = (f64["Uw"].cpu().numpy(), f64["Pw"].cpu().numpy())
    nn = mix["nn"]
    U64, P64, Um, Pm = (f64["U"].cpu().numpy(), f64["P"].cpu().numpy(), mix["U"].cpu().numpy(),
                        mix["P"].cpu().numpy())

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    # dominant kernel: momentum operator apply (BiCGSTAB s = r - alpha q on the fly, t = A s, 3 dots)
    ne = N_EL * N_EL
    reps = 50
    ms_m = mix["ctx"].bench_momentum_apply(reps)
    ms_d = f64["ctx"].bench_momentum_apply(reps)
    bytes_m = 16 * ne * 4 + 6 * nn * 8      # FP32 Ce + r, q read, t written (2 components, FP64)
    bytes_d = 16 * ne * 8 + 6 * nn * 8
    ach_m = bytes_m / (ms_m * 1e-3) / 1e9
    ach_d = bytes_d / (ms_d * 1e-3) / 1e9

    copy_gbs = copy_bandwidth()
    pois64 = poisson_alone("fp64")
    poism = poisson_alone("mixed")
    tg64 = taylor_green_64(not args.no_cpu)
    grad = None if args.no_grad else gradient_256()
    ch3d = None if args.no_3d else channel_3d(args.steps_3d)
    ch3d_q2 = None if args.no_3d else channel_3d_q2()
    cpu = cpu1 = None
    if not args.no_cpu:
        cpu = cpu_baseline(*cpu_start, args.warmup + 1, args.cpu_steps)
        cpu1 = cpu_baseline(*cpu_start, args.warmup + 1, max(1, args.cpu_steps - 1), threads=1)

    steps_s_m = args.steps / mix["sec"]
    steps_s_d = args.steps / f64["sec"]
    traffic = _traffic_from_profile()
    line = {
        "metric": "timesteps/s 1024² Q1: mixed-prec vs FP64 MG-PCG at 1e-8 residual",
        "value": steps_s_m,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mix["sec"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (mixed)",
        "data": "synthetic (seeded N(0,1e-3) initial velocity)",
        "config": {"workload": "lid-driven cavity 1024x1024 Q1, Re=1000, dt=5e-4, BiCGSTAB + MG-PCG "
                               "(1024->16, pinned p), tol 1e-8, Ce FP32, native step loop (fcfd_four_step_rollout)",
                   "elements": ne, "nodes": nn, "mg_levels": 7},
        "fp64": {"timesteps_per_s": steps_s_d, "ms_per_step": 1e3 * f64["sec"] / args.steps,
                 "mean_iters": [float(np.mean([i[0] for i in f64["its"]])),
                                float(np.mean([i[1] for i in f64["its"]]))],
                 "python_call_per_step_timesteps_per_s": 1.0 / f64["py_step_s"]},
        "mixed": {"mean_iters": [float(np.mean([i[0] for i in mix["its"]])),
                                 float(np.mean([i[1] for i in mix["its"]]))],
                  "python_call_per_step_timesteps_per_s": 1.0 / mix["py_step_s"]},
        "speedup_mixed_vs_fp64": steps_s_m / steps_s_d,
        "mixed_vs_fp64_rel_l2": {"u": rel(Um[:nn], U64[:nn]), "v": rel(Um[nn:], U64[nn:]), "p": rel(Pm, P64)},
        "poisson_1024_neumann": {"fp64": pois64, "mixed": poism},
        "taylor_green_64": tg64,
        "gradient_256": grad,
        "channel_3d": ch3d,
        "channel_3d_q2": ch3d_q2,
        "roofline": {"bound": "hbm", "achieved": ach_m, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_m / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_mom_k2t<CeF32, 64, 4, 1> (BiCGSTAB s/t: tiled momentum operator apply + 3 dots)",
                     "bytes_per_launch": bytes_m, "us_per_launch": ms_m * 1e3,
                     "device_copy_gbs": copy_gbs, "frac_of_device_copy": ach_m / copy_gbs,
                     "fp64_ce": {"achieved": ach_d, "bytes_per_launch": bytes_d, "us_per_launch": ms_d * 1e3}},
        "cpu_baseline": None if cpu is None else (
            {"error": cpu["error"]} if "error" in cpu else
            {"value": cpu["value"], "unit": "timesteps/s", "cores": cpu["threads"], "kind": "port",
             "cores_available": cpu["cores"],
             "sample": f"steps {cpu['window'][0]}-{cpu['window'][1]} of the same 1024^2 cavity (inside the timed "
                       f"window of `value`), FP64 oracle (NumPy), started from the FP64 GPU state after step "
                       f"{args.warmup} and one untimed oracle step; OMP/OpenBLAS/MKL threads = {cpu['threads']} "
                       f"(NumPy's gathers and scatters are single-threaded); iters {cpu['its']}",
             "one_thread": None if cpu1 is None or "error" in cpu1 else
             {"value": cpu1["value"], "steps": cpu1["steps"], "iters": cpu1["its"]}}),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
