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


def timestep_loop(t_final, dt, save_interval, print_interval, compute_func, timestep_func, parameters,
                  dirichlet_vals, initial_states, S_func=None, phi_S_func=None, save2file=None, h5name=None):
    """simulation.py:229-360: per step compute -> timestep, iteration counts and
    solutions recorded on save steps, progress printed on steps <= 5, every
    print_interval and the last step.  ``save2file`` needs h5py (not installed
    in this image) and raises if requested."""
    if save2file:
        raise NotImplementedError("HDF5 trajectory streaming needs h5py, which is not available")
    t = nat.torch()
    has_phi = len(initial_states) == 3
    U, P = nat.as_dev(initial_states[0]), nat.as_dev(initial_states[1])
    phi = nat.as_dev(initial_states[2]) if has_phi else None
    U_d = dirichlet_vals[0]
    phi_d = dirichlet_vals[1] if has_phi else None
    U_iter_list, P_iter_list, step_list = [], [], []
    phi_iter_list, phi_list = [], []
    F_p, F = compute_func(U_d, S_func, 0, 0)
    compile_start = time.perf_counter()
    if has_phi:
        _ = timestep_func(U, P, phi, dt, *parameters, U_d, phi_d, F_p, F, phi_S_func)
    else:
        _ = timestep_func(U, P, dt, *parameters, U_d, F_p, F)
    compile_time = time.perf_counter() - compile_start
    print(f"Run to compile took {compile_time:.2f} s\n")
    num_steps = int(t_final / dt)
    if num_steps * dt < t_final:
        num_steps += 1
    print("num steps:", num_steps)
    U_list = [nat.to_numpy(U)]
    if has_phi:
        phi_list = [nat.to_numpy(phi)]
    total_start = time.perf_counter()
    for step in range(1, num_steps + 1):
        U_n, P_n = U, P
        phi_n = phi
        t_n = (step - 1) * dt
        tt = step * dt
        start = time.perf_counter()
        F_p, F = compute_func(U_d, S_func, t_n, tt)
        if has_phi:
            U, P, phi, U_iters, P_iters, phi_iters = timestep_func(U_n, P_n, phi_n, dt, *parameters, U_d, phi_d,
                                                                   F_p, F, phi_S_func)
        else:
            U, P, U_iters, P_iters = timestep_func(U_n, P_n, dt, *parameters, U_d, F_p, F)
        end = time.perf_counter()
        if step % save_interval == 0 or step == num_steps:
            U_list.append(nat.to_numpy(U))
            U_iter_list.append(U_iters)
            P_iter_list.append(P_iters)
            step_list.append(step)
            if has_phi:
                phi_list.append(nat.to_numpy(phi))
                phi_iter_list.append(phi_iters)
        if step <= 5 or step % print_interval == 0 or step == num_steps:
            diff = float(t.linalg.norm(U - U_n) / t.linalg.norm(U))
            if has_phi:
                print(f"End step {step}, time {tt:.2f}, took {((end - start) * 1000):.2f} ms, {U_iters} U iters, "
                      f"{P_iters} P iters, {phi_iters} phi iters, diff = {diff:.6f}")
            else:
                print(f"End step {step}, time {tt:.2f}, took {((end - start) * 1000):.2f} ms, {U_iters} U iters, "
                      f"{P_iters} P iters, diff = {diff:.6f}")
    total_time = time.perf_counter() - total_start
    print(f"\nTotal timestepping time: {total_time:.2f} s")
    print(f"Avg of {((total_time / num_steps) * 1000):.2f} ms per timestep")
    print(f"Avg of {(((total_time + compile_time) / num_steps) * 1000):.2f} ms per timestep including compile time\n")
    U_iter_list = np.array(U_iter_list)
    P_iter_list = np.array(P_iter_list)
    step_list = np.array(step_list)
    U_list = np.stack(U_list, axis=0)
    if has_phi:
        return U, P, phi, U_iter_list, P_iter_list, np.array(phi_iter_list), step_list, U_list, np.stack(phi_list)
    return U, P, U_iter_list, P_iter_list, step_list, U_list


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
