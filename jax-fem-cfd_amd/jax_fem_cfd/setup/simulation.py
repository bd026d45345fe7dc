"""Configuration, set-up and the host time-stepping loop.

Same dataclasses, ``setup`` return tuple, ``set_face_velocity`` and
``timestep_loop`` signature as jax_fem_cfd/setup/simulation.py.  Extension:
``StandardSolverConfig`` takes two optional trailing fields, ``precision``
("fp64": the reference's algorithm in FP64; "mixed": FP32 multigrid inside an
FP64 flexible CG and reduced-precision stored convection matrices) and
``ce_format`` ("f64", "f32", "f16s", "bf16s"; default follows precision).
"""
import math
import time
from dataclasses import dataclass
from typing import Dict, List, Literal, Optional, Union

import numpy as np

from .. import _native as nat


@dataclass
class SimpleMeshConfig:
    num_elem: List[int]
    domain_size: List[float]
    shape_func_ord: int
    periodic: bool
    partial_dirichlet: Dict = None


@dataclass
class StandardSolverConfig:
    dirichlet_faces: List[int]
    symmetry_faces: List[int]
    outlet_faces: List[int]
    streamline_diff: bool
    solver_tol: float
    set_zeroP: bool
    pressure_precond: Optional[str] = None
    final_multigrid_mesh: Optional[List[int]] = None
    precision: str = "fp64"
    ce_format: Optional[str] = None


@dataclass
class ADEConfig:
    phi_dirichlet_faces: List[int]
    phi_SUPG: bool


@dataclass
class PassiveTransportConfig:
    NS_config: StandardSolverConfig
    ADE_config: ADEConfig


@dataclass
class Config:
    mesh: Literal["simple"]
    solver: Literal["standard", "passive_transport"]
    mesh_config: SimpleMeshConfig
    solver_config: Union[StandardSolverConfig, PassiveTransportConfig]

    def __post_init__(self):
        self.validate()

    def validate(self):
        """simulation.py:51-93 (same ValueErrors), plus the precision fields."""
        if self.mesh == "simple":
            if not isinstance(self.mesh_config, SimpleMeshConfig):
                raise ValueError("Mesh type and config type must match")
        else:
            raise ValueError(f"Unknown mesh type: {self.mesh}")
        if self.solver == "standard":
            if not isinstance(self.solver_config, StandardSolverConfig):
                raise ValueError("Solver type and config type must match")
        elif self.solver == "passive_transport":
            if not isinstance(self.solver_config, PassiveTransportConfig):
                raise ValueError("Solver type and config type must match")
        else:
            raise ValueError(f"Unknown solver type: {self.solver}")
        if self.solver == "standard":
            sc = self.solver_config
            if len(self.mesh_config.num_elem) == 3:
                # hex27 is built here; the reference raises "3D quadratic shape functions not yet
                # implemented" (simulation.py:73-74)
                # 3D multigrid is built here (the box-grid V-cycle); the reference raises
                # "3D multigrid not yet implemented" (simulation.py:75-76)
                if self.mesh_config.periodic:
                    raise ValueError("3D periodic boundary conditions not yet implemented")
            if sc.pressure_precond == "multigrid":
                num_elemf = sc.final_multigrid_mesh
                if num_elemf is None:
                    raise ValueError("Must specify final_multigrid_mesh to use multigrid")
                for n, f in zip(self.mesh_config.num_elem, num_elemf):
                    if f > n:
                        raise ValueError("final_multigrid_mesh is too large")
                    r = math.log2(n // f)
                    if not math.isclose(r, round(r)):
                        raise ValueError(f"Invalid final_multigrid_mesh: log2({n}/{f}) is not an integer")
            if sc.precision not in nat.PREC:
                raise ValueError(f"Unknown precision: {sc.precision}")
            if sc.ce_format is not None and sc.ce_format not in nat.CE_FMT:
                raise ValueError(f"Unknown ce_format: {sc.ce_format}")


def simple_mesh(mesh_config: SimpleMeshConfig, NS_config: StandardSolverConfig):
    """simulation.py:96-183: mesh + BC index sets (host set-up arrays)."""
    from . import regular_mesh as rm
    domain_size = mesh_config.domain_size
    num_elem = mesh_config.num_elem
    ndim = len(num_elem)
    if mesh_config.shape_func_ord == 1:
        if ndim == 2:
            mesh_func = rm.create_periodic_mesh_2d if mesh_config.periodic else rm.create_mesh_2d
        else:
            mesh_func = rm.create_mesh_3d
    else:
        if ndim == 2:
            mesh_func = rm.create_periodic_mesh_2d_quad9 if mesh_config.periodic else rm.create_mesh_2d_quad9
        else:
            mesh_func = rm.create_mesh_3d_hex27
    node_coords, nconn, surfnodes, bconns, nvecs = mesh_func(*num_elem, *domain_size)
    nn = node_coords.shape[0]
    h = mesh_config.domain_size[0] / mesh_config.num_elem[0]
    bconn, nvec, gamma_d, U_d, gamma_p = None, None, None, None, None
    dirichlet_nodes, outlet_nodes, bconn_list, nvec_list = [], [], [], []
    if surfnodes is not None:
        surfnodes = list(surfnodes)
    for face in NS_config.dirichlet_faces or []:
        dirichlet_nodes.append(surfnodes[face])
        bconn_list.append(bconns[face])
        nvec_list.append(nvecs[face])
    for face in NS_config.outlet_faces or []:
        outlet_nodes.append(surfnodes[face])
    if mesh_config.partial_dirichlet:
        for face, (start_frac, end_frac, remainder_type) in mesh_config.partial_dirichlet.items():
            face_nodes = surfnodes[face]
            num_face_elem = bconns[face].shape[0]
            s_idx = round(start_frac * num_face_elem)
            e_idx = round(end_frac * num_face_elem)
            bconn_list.append(bconns[face][s_idx:e_idx])
            nvec_list.append(nvecs[face][s_idx:e_idx])
            if mesh_config.shape_func_ord == 2:
                s_idx, e_idx = 2 * s_idx, 2 * e_idx
            surfnodes[face] = face_nodes[s_idx:e_idx + 1]
            dirichlet_nodes.append(face_nodes[s_idx:e_idx + 1])
            if remainder_type == "outlet":
                outlet_nodes.append(face_nodes[:s_idx])
                outlet_nodes.append(face_nodes[e_idx + 1:])
    if dirichlet_nodes:
        g = np.unique(np.concatenate(dirichlet_nodes))
        gamma_d = np.concatenate([g + nn * i for i in range(ndim)]).astype(np.int64)
        U_d = np.zeros(gamma_d.shape[0])
        bconn = np.concatenate(bconn_list, axis=0)
        nvec = np.concatenate(nvec_list, axis=0)
    if outlet_nodes:
        gamma_p = np.unique(np.concatenate(outlet_nodes)).astype(np.int64)
    else:
        gamma_p = 0
    return node_coords, nconn, surfnodes, nn, h, bconn, nvec, gamma_d, U_d, gamma_p, mesh_func


def setup(config: Config):
    """simulation.py:185-217."""
    if config.solver == "standard":
        mesh_data = simple_mesh(config.mesh_config, config.solver_config)
    else:
        mesh_data = simple_mesh(config.mesh_config, config.solver_config.NS_config)
    node_coords, nconn, surfnodes, nn, h, bconn, nvec, gamma_d, U_d, gamma_p, mesh_func = mesh_data
    if config.solver == "standard":
        from ..solvers.standard import setup_solver
        compute, timestep = setup_solver(node_coords, nconn, bconn, nvec, gamma_d, gamma_p, nn, mesh_func, config)
        return compute, timestep, node_coords, surfnodes, h, nn, gamma_d, U_d
    from ..solvers.passive_transport import setup_solver
    mc = config.mesh_config
    nnx, nny = (mc.num_elem[0] + 1, mc.num_elem[1] + 1) if mc.shape_func_ord == 1 else \
        (2 * mc.num_elem[0] + 1, 2 * mc.num_elem[1] + 1)
    faces = config.solver_config.ADE_config.phi_dirichlet_faces
    gamma_d_phi = None if not faces else np.sort(np.unique(np.concatenate([surfnodes[i] for i in faces])))
    compute, timestep = setup_solver(node_coords, nconn, bconn, nvec, gamma_d, gamma_p, nn, gamma_d_phi, mesh_func,
                                     config)
    return compute, timestep, node_coords, surfnodes, h, nn, nnx, nny, gamma_d, U_d, gamma_d_phi


def set_face_velocity(U_d, nodes, gamma_d, values, nn, var):
    """simulation.py:219-226: searchsorted into the sorted gamma_d."""
    idx = np.searchsorted(nat.to_numpy(gamma_d), nat.to_numpy(nodes) + nn * (var - 1))
    out = np.array(nat.to_numpy(U_d), dtype=float, copy=True)
    out[idx] = values
    return out


def _n_steps(t_final, dt):
    """Steps to reach t_final: ceil(t_final / dt), with t_final / dt truncated first (the
    reference's count at simulation.py:258-260)."""
    n = int(t_final / dt)
    return n + 1 if n * dt < t_final else n


class _Record:
    """What timestep_loop keeps: the initial and every saved state, the iteration counts and the
    step index of every save."""

    def __init__(self, U0, phi0):
        self.U = [nat.to_numpy(U0)]
        self.phi = None if phi0 is None else [nat.to_numpy(phi0)]
        self.its = []
        self.steps = []

    def save(self, step, U, its, phi=None):
        self.U.append(nat.to_numpy(U))
        if self.phi is not None:
            self.phi.append(nat.to_numpy(phi))
        self.its.append(its)
        self.steps.append(step)

    def counts(self, k):
        return np.array([c[k] for c in self.its])


def timestep_loop(t_final, dt, save_interval, print_interval, compute_func, timestep_func, parameters,
                  dirichlet_vals, initial_states, S_func=None, phi_S_func=None, save2file=None, h5name=None):
    """The reference's host time loop (simulation.py:229-360) over the native step.

    One untimed warm-up call of the step first (the reference's compile run; its result is
    dropped), then ``_n_steps(t_final, dt)`` steps, each calling ``compute_func(U_d, S_func,
    t_n, t_n + dt)`` and then the step.  The state, its iteration counts and the step index are
    kept at every ``save_interval``-th step and at the last one (the initial state is the first
    entry of the state list).  A progress line is printed for the first five steps, every
    ``print_interval``-th step and the last one.

    Returns (U, P, U_iters, P_iters, steps, U_list), or with a passive scalar (three initial
    states) (U, P, phi, U_iters, P_iters, phi_iters, steps, U_list, phi_list).  ``save2file``
    (HDF5 streaming) needs h5py, which this image does not have, and raises if requested."""
    if save2file:
        raise NotImplementedError("HDF5 trajectory streaming needs h5py, which is not available")
    t = nat.torch()
    passive = len(initial_states) == 3
    U, P = nat.as_dev(initial_states[0]), nat.as_dev(initial_states[1])
    phi = nat.as_dev(initial_states[2]) if passive else None
    U_d, phi_d = dirichlet_vals[0], (dirichlet_vals[1] if passive else None)

    def step_once(U_n, P_n, phi_n, F_p, F):
        if passive:
            return timestep_func(U_n, P_n, phi_n, dt, *parameters, U_d, phi_d, F_p, F, phi_S_func)
        U1, P1, it_u, it_p = timestep_func(U_n, P_n, dt, *parameters, U_d, F_p, F)
        return U1, P1, None, it_u, it_p, None

    w0 = time.perf_counter()
    step_once(U, P, phi, *compute_func(U_d, S_func, 0, 0))
    t_warm = time.perf_counter() - w0
    n = _n_steps(t_final, dt)
    print(f"warm-up step (graph capture, first launches): {t_warm:.2f} s; running {n} steps")
    rec = _Record(U, phi)
    loop0 = time.perf_counter()
    for k in range(1, n + 1):
        s0 = time.perf_counter()
        U_prev = U
        F_p, F = compute_func(U_d, S_func, (k - 1) * dt, k * dt)
        U, P, phi, it_u, it_p, it_phi = step_once(U, P, phi, F_p, F)
        ms = (time.perf_counter() - s0) * 1e3
        its = (it_u, it_p) if not passive else (it_u, it_p, it_phi)
        if k % save_interval == 0 or k == n:
            rec.save(k, U, its, phi)
        if k <= 5 or k % print_interval == 0 or k == n:
            change = float(t.linalg.norm(U - U_prev) / t.linalg.norm(U))
            extra = f", scalar {it_phi}" if passive else ""
            print(f"step {k}/{n} (t = {k * dt:.4g}): {ms:.2f} ms; iterations momentum {it_u}, pressure {it_p}"
                  f"{extra}; |U - U_prev| / |U| = {change:.3e}")
    total = time.perf_counter() - loop0
    print(f"{n} steps in {total:.2f} s: {total / n * 1e3:.2f} ms per step "
          f"({(total + t_warm) / n * 1e3:.2f} ms per step with the warm-up call)")
    steps = np.array(rec.steps)
    U_list = np.stack(rec.U, axis=0)
    if passive:
        return (U, P, phi, rec.counts(0), rec.counts(1), rec.counts(2), steps, U_list,
                np.stack(rec.phi, axis=0))
    return U, P, rec.counts(0), rec.counts(1), steps, U_list


def timestep_loop_scan(dt, num_steps, compute_func, timestep_func, rho, mu, D, U_d, phi_d, U0, P0, phi0, S_func,
                       checkpoint=20):
    """simulation.py:362-374: F_p, F = compute(U_d, S_func, 0, 0) once, then num_steps
    passive-scalar steps from (U0, P0, phi0); returns the final (U, P, phi).

    Differentiable w.r.t. U0, P0, phi0 (float64 CUDA tensors with requires_grad) through
    torch autograd; the backward pass is the hand-written adjoint of each step
    (fcfd_passive_timestep_vjp) swept last step first, with the states recomputed from
    checkpoints kept every `checkpoint` steps.

    The reference's scan body hands F_p and phi_d to the step in each other's positions
    (simulation.py:368); here each goes to its own parameter.  Both agree whenever the
    reference's call runs at all (numeric phi_d, no scalar Dirichlet nodes, F_p = 0)."""
    from ..solvers.adjoint import rollout
    F_p, F = compute_func(U_d, S_func, 0, 0)
    info = getattr(timestep_func, "rollout_info", None)
    if info is None:
        raise TypeError("timestep_func must be the passive-scalar timestep returned by setup()")
    ctx = timestep_func.context
    return rollout(ctx, int(num_steps), U0, P0, phi0, dt, rho, mu, D,
                   U_d if info["has_dirichlet"] else None, phi_d if info["has_phi_dirichlet"] else None,
                   F_p, F, info["supg"], checkpoint)
