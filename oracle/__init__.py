"""CPU FP64 restatement of the jax-fem-cfd matrix-free Navier-Stokes timestep.

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP product
path: it may be imported by ``tests/``, by ``__graft_entry__.smoke()`` and by
the ``cpu_baseline`` leg of ``bench.py`` -- nowhere else.  The product package
(``jax-fem-cfd_amd/``) never imports it and has no CPU fallback.

Every function is written from the mathematics and cites the reference
function (``path:line`` relative to the reference repository) whose
behaviour it restates.  Arrays are NumPy, FP64 throughout, integers int64.

Modules
-------
mesh        structured Q1 meshes and boundary index sets
elements    shape functions, element matrices, convection matrices,
            boundary integral
emult       element-by-element matvecs (gather -> dense -> scatter)
solvers     CG / BiCGSTAB with the reference's exact control flow
multigrid   geometric V-cycle preconditioner
ns          momentum / pressure solves, 4-step projection, host loop,
            passive scalar transport

Pinning (what checks each function; see tests/test_oracle_*.py):
  mesh        closed-form index formulas, incidence counts
  elements    closed forms (Me, Le, Ge), row-sum and symmetry invariants,
              high-order quadrature brute force for Ce, periodic
              skew-symmetry of Ce(beta=1/2)
  emult       scipy.sparse assembly of the same element matrices,
              adjoint identity <Gp,u> = <p,Du>, 9-point stencil
  solvers     scipy.sparse.linalg.spsolve solution of the assembled
              system, true-residual bound
  multigrid   R = I^T on non-periodic grids, I 1 = 1, R 1 row weights,
              V-cycle linearity (Neumann form), dense two-level restatement;
              3D: P3 against kron products of 1D prolongations, R3 = P3^T,
              trilinear reproduction, per-level Le closed forms, dense
              two-level restatement, PCG iteration counts
  ns          Taylor-Green analytic decay, discrete divergence reduction,
              Ghia cavity profile (coarse)
"""
