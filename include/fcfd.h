/*
 * fcfd.h -- C ABI of the MI355X (gfx950) matrix-free Navier-Stokes timestep.
 *
 * Drop-in boundary for the hot path of jax-fem-cfd: the 4-step pressure
 * projection (per-element convection matrices, element-by-element matvecs,
 * BiCGSTAB momentum solve, multigrid-preconditioned CG pressure solve).
 * The reference has no FFI of its own -- its boundary is the Python call
 * surface -- so each entry point below names the reference function it
 * replaces (paths relative to the reference repository).
 *
 * Conventions
 *   - All vector arguments are DEVICE pointers (hipMalloc / torch CUDA
 *     tensors), FP64, in the reference layout: nodal vectors (nn,), velocity
 *     component-major [u; v] (2*nn,), element matrices (ne, 4, 4) row-major.
 *   - Index arrays passed to fcfd_ctx_create are HOST pointers (int64).
 *   - Every call is asynchronous on the context stream unless it returns
 *     iteration counts (those synchronise the stream once at the end).
 *   - Return value: FCFD_OK (0) or a negative FCFD_ERR_*; fcfd_last_error()
 *     gives the message.  Solver iteration counts mirror the reference:
 *     BiCGSTAB returns -10 (rho = 0) or -11 (omega or alpha = 0) on
 *     breakdown; *converged tells maxiter exhaustion from success.
 *   - A context is not thread safe; one stream per context.
 */
#ifndef FCFD_H
#define FCFD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fcfd_ctx fcfd_ctx;

enum {
    FCFD_OK = 0,
    FCFD_ERR_ARG = -1,
    FCFD_ERR_HIP = -2,
    FCFD_ERR_UNSUPPORTED = -3,
    FCFD_ERR_STATE = -4
};

/* pressure preconditioner (StandardSolverConfig.pressure_precond) */
enum { FCFD_PC_NONE = 0, FCFD_PC_JACOBI = 1, FCFD_PC_APPROX_INV = 2, FCFD_PC_MULTIGRID = 3 };

/* solver precision policy */
enum {
    FCFD_PREC_FP64 = 0,   /* the reference's algorithm, FP64 throughout            */
    FCFD_PREC_MIXED = 1   /* FP64 outer flexible CG / FP64 Krylov vectors, FP32
                             multigrid, reduced-precision stored element matrices  */
};

/* storage format of the per-element convection matrices */
enum { FCFD_CE_F64 = 0, FCFD_CE_F32 = 1, FCFD_CE_F16S = 2, FCFD_CE_BF16S = 3 };

typedef struct fcfd_mesh_desc {
    int32_t ndim;          /* 2, or 3 for hex8 boxes (create_mesh_3d, regular_mesh.py:45-105) */
    int32_t nex, ney;      /* elements per direction */
    double lx, ly;         /* domain size */
    int32_t periodic;      /* 1: create_periodic_mesh_2d, 0: create_mesh_2d (3D: must be 0) */
    int32_t nez;           /* ndim = 3: elements in z */
    double lz;             /* ndim = 3: domain size in z */
    int32_t order;         /* ndim = 2: 1 (or 0) quad4, 2 quad9 (create_mesh_2d_quad9,
                              regular_mesh.py:107-145, :182-218); vectors then live on the
                              (2 nex + 1) x (2 ney + 1) node grid (2 nex x 2 ney periodic).
                              ndim = 3: 1 (or 0) hex8, 2 hex27 (no counterpart in the
                              reference, simulation.py:73-74 rejects it): vectors on the
                              (2 nex + 1) x (2 ney + 1) x (2 nez + 1) node grid */
} fcfd_mesh_desc;

typedef struct fcfd_solver_desc {
    double tol;                 /* solver_tol: BiCGSTAB, PCG and the coarse CG   */
    int32_t set_zeroP;          /* pin gamma_p (1) or pure-Neumann mean form (0) */
    int32_t precond;            /* FCFD_PC_*                                     */
    int32_t mg_levels;          /* multigrid levels incl. the fine one (>= 1)    */
    int32_t precision;          /* FCFD_PREC_*                                   */
    int32_t ce_format;          /* FCFD_CE_* used by the timestep                */
    int32_t maxiter_momentum;   /* <= 0: 10 * size, as the reference             */
    int32_t maxiter_pressure;   /* <= 0: 1000, as the reference                  */
    const int64_t* gamma_d;     /* sorted velocity Dirichlet DOF ids or NULL     */
    int64_t n_gamma_d;
    const int64_t* gamma_p;     /* fine-level pressure pins (pinned form)        */
    int64_t n_gamma_p;
    const int64_t* mg_gamma_p;      /* per-level pins, concatenated              */
    const int64_t* mg_gamma_p_off;  /* mg_levels + 1 offsets into mg_gamma_p     */
    const int64_t* bconn;       /* (n_bedge, 2) Dirichlet boundary edges; quad9: */
                                /* (n_bedge, 3); hex8: (n_bedge, 4) faces;       */
                                /* hex27: (n_bedge, 9) quad9 faces               */
    const int64_t* nvec;        /* (n_bedge, ndim) outward normals               */
    int64_t n_bedge;
    const int64_t* gamma_d_phi; /* passive-scalar Dirichlet nodes or NULL         */
    int64_t n_gamma_d_phi;
    int32_t approx_inv_variant; /* approx_inv closure: 0 that of standard.py:225-234 (gamma_p rows of
                                   both inner products 0), 1 that of passive_transport.py:154-162
                                   (those rows keep D1x and D1L) -- quad4 contexts */
} fcfd_solver_desc;

/* ---- context --------------------------------------------------------------------------- */

/* replaces setup_solver (jax_fem_cfd/solvers/standard.py:191-247): element
 * constants from element 0, lumped mass, multigrid hierarchy
 * (jax_fem_cfd/core/multigrid.py:9-88), Dirichlet masks; allocates every
 * workspace the hot calls use.  stream: hipStream_t (NULL = default). */
int fcfd_ctx_create(const fcfd_mesh_desc* mesh, const fcfd_solver_desc* solver, void* stream,
                    fcfd_ctx** out);
int fcfd_ctx_destroy(fcfd_ctx* ctx);
int fcfd_set_stream(fcfd_ctx* ctx, void* stream);
const char* fcfd_last_error(void);
/* nn, ne, number of multigrid levels, nodes of level l (l = -1: returns levels) */
int fcfd_ctx_info(const fcfd_ctx* ctx, int64_t* nn, int64_t* ne, int32_t* levels);
int fcfd_level_nodes(const fcfd_ctx* ctx, int32_t level, int64_t* nn, int32_t* nnx, int32_t* nny);

/* ---- mesh --------------------------------------------------------------------------- */

/* node coordinates (nn, 2) and connectivity (ne, 4) from the device index
 * math the kernels use.  replaces create_mesh_2d / create_periodic_mesh_2d
 * (jax_fem_cfd/setup/regular_mesh.py:5-43, :149-180). */
int fcfd_mesh_fill(fcfd_ctx* ctx, double* coords, int64_t* nconn);

/* host copies of the FP64 element constants (16 doubles each, row-major):
 * which = 0 Me, 1 Le, 2 Ge[0], 3 Ge[1], 4 Le_off; plus lumped mass (2*nn, device).
 * replaces lumped_mass_matrix / laplacian_element_calc / gradient_element_calc
 * (jax_fem_cfd/core/element_calc.py:9-40). */
int fcfd_element_constant(const fcfd_ctx* ctx, int32_t which, double* out16);
/* 3D contexts: the hex8 constants (64 doubles each, row-major): which = 0 Me, 1 Le, 2..4 Ge[0..2].
 * replaces lumped_mass_matrix / laplacian_element_calc / gradient_element_calc on hex8
 * (element_calc.py:9-40, shape_functions.py:47-86). */
int fcfd_element_constant3(const fcfd_ctx* ctx, int32_t which, double* out64);
/* y = M_c x for every velocity component (ndim * nn, device): the momentum source term of
 * compute (standard.py:184-186).  replaces consistent_mass_mult (element_mult.py:11-36; its
 * ndim = 3 branch fails in the reference, here it is the per-component product). */
int fcfd_mass_matvec(fcfd_ctx* ctx, const double* x, double* y);
/* Element matrix of any context, n = nen * nen doubles row-major (quad4 16, quad9 81, hex8 64,
 * hex27 729): which = 0 Me, 1 Le, 2..(1 + ndim) Ge[d], 5 Le_off (2D, hex27).  replaces lumped_mass_matrix /
 * laplacian_element_calc / gradient_element_calc (element_calc.py:9-40) on the context's
 * element (quad4 / quad9 shape_functions.py:27-45, :108-159; hex8 :47-86). */
int fcfd_element_matrix(const fcfd_ctx* ctx, int32_t which, double* out, int32_t n);
int fcfd_lumped_mass(fcfd_ctx* ctx, double* M);
/* L_diag of level l (device out).  replaces laplacian_diag (element_calc.py:99-109) */
int fcfd_laplacian_diag(fcfd_ctx* ctx, int32_t level, double* out);

/* ---- element calculations ----------------------------------------------------------- */

/* per-element convection matrices (ne, 4, 4) in the reference layout.
 * replaces conservative_convec_elem_calc (jax_fem_cfd/core/element_calc.py:165-185) */
int fcfd_conv_elem_calc(fcfd_ctx* ctx, const double* U, double beta, double* Ce);

/* F_p = int_Gamma q (U_d . n).  replaces u_dot_n_boundary_integral_1d
 * (element_calc.py:111-135) as called by compute (standard.py:167-188).
 * U_d: device (n_gamma_d,). */
int fcfd_boundary_integral(fcfd_ctx* ctx, const double* U_d, double* F_p);

/* ---- element-by-element matvecs (jax_fem_cfd/core/element_mult.py) ------------------ */

/* y = A x per component with one shared element matrix A (16 doubles, host).
 * Covers consistent_mass_mult (:11-36), laplacian_element_mult (:38-49,
 * also with Le_off), gradient_element_mult (:51-77, one call per G_i) and
 * divergence_element_mult (:133-157, A = G_i^T, accumulate = 1).
 * level selects the multigrid grid (0 = fine).  accumulate: y += A x. */
int fcfd_elem_mult_const(fcfd_ctx* ctx, int32_t level, const double* A16, const double* x,
                         double* y, int32_t accumulate);

/* y_c = sum_e (K0 + c1 A1_e + c2 A2_e) x_c per component (ncomp 1..3);
 * K0: shared element matrix (16 doubles, host, NULL = 0); A1, A2: per-element
 * matrices (ne,4,4) device or NULL.  Replaces visc_conv_mult (:79-104;
 * K0 = mu Le, A1 = Ce, c1 = rho), scalar_visc_conv_mult (:159-172),
 * scalar_stabilized_visc_conv_mult (:174-187; A2 = Se) and
 * scalar_consistent_mass_SUPG_mult (:190-201; K0 = Me, A1 = Me_tau). */
int fcfd_elem_mult_stored(fcfd_ctx* ctx, const double* K0_16, const double* A1, double c1,
                          const double* A2, double c2, int32_t ncomp, const double* x, double* y);

/* ---- multigrid and solvers --------------------------------------------------------- */

/* z = V(r), one V-cycle from zero.  replaces v_cycle (jax_fem_cfd/core/multigrid.py:231-330);
 * quad4, quad9 and (new: the reference has none, simulation.py:75-76) hex8 box contexts */
int fcfd_vcycle(fcfd_ctx* ctx, const double* r, double* z);

/* interpolation_operator (multigrid.py:90-140): bilinear prolongation of xH on multigrid level
 * level + 1 of the context to xh on level `level` (periodic grids wrap).  2D quad4 contexts. */
int fcfd_interpolate(fcfd_ctx* ctx, int32_t level, const double* xH, double* xh);

/* restriction_operator (multigrid.py:142-212): R = I^T without the 1/4 factor, xh on level
 * `level` to xH on level + 1; periodic grids restrict the padded grid and drop its wrap row and
 * column, as the reference does.  2D quad4 contexts. */
int fcfd_restrict(fcfd_ctx* ctx, int32_t level, const double* xh, double* xH);

/* single_jacobi_iter (multigrid.py:214-228): one damped Jacobi sweep (omega = 2/3) on level
 * `level`: x_out = omega (rhs - Le_off x) / L_diag + (1 - omega) x.  Le_off16: host double[16];
 * L_diag: device (nn of the level); gamma_p: device int64[n_gamma_p].  set_zeroP: x zeroed at
 * gamma_p before the sweep and x_out after it; otherwise the mean of x_out is subtracted.
 * x_out must not alias x.  2D quad4 contexts. */
int fcfd_jacobi_sweep(fcfd_ctx* ctx, int32_t level, const double* Le_off16, const double* x, const double* rhs,
                      const double* L_diag, const int64_t* gamma_p, int64_t n_gamma_p, int32_t set_zeroP,
                      double* x_out);

/* PCG on the (pinned or Neumann) pressure Laplacian with the context's
 * preconditioner; x holds x0 on entry.  Mirrors the cg call inside
 * solve_pressure (standard.py:114-117) and _cg_solve
 * (jax_fem_cfd/utils/jax_iterative_solvers.py:87-120): stops on the recursive
 * residual r.r <= max(tol^2 b.b, 0) or maxiter. */
int fcfd_pressure_cg(fcfd_ctx* ctx, const double* b, double* x, double tol, int32_t maxiter,
                     int32_t* iters, int32_t* converged);

/* solve_pressure (standard.py:85-125), gamma = 0: rhs = rho/dt (D U* - F_p),
 * pin or mean-project, x0 = P_n, PCG, post-process.  F_p may be NULL (zero). */
int fcfd_solve_pressure(fcfd_ctx* ctx, const double* U_star, const double* P_n,
                        const double* F_p, double rho, double dt, double* P, int32_t* iters);

/* solve_momentum (standard.py:14-43), theta = 1/2, gamma = 1: BiCGSTAB
 * (jax_iterative_solvers.py:123-167) on (rho/dt) M_c + theta (rho Ce + mu Le)
 * with Dirichlet rows -> identity.  Ce: (ne,4,4) device (NULL: the context
 * builds it from U_n with beta = 1/2).  F may be NULL. */
int fcfd_solve_momentum(fcfd_ctx* ctx, const double* U_n, const double* P_n, const double* Ce,
                        double rho, double mu, double dt, const double* U_d, const double* F,
                        double* U_hat, int32_t* iters);

/* four_step_timestep (standard.py:148-165): Ce(U_n), momentum, U*, pressure,
 * projection, Dirichlet values.  Uses the context's precision policy.
 * U_d: device (n_gamma_d,) or NULL; F_p, F: device or NULL (zero).
 * iters: host int32[2] = (U_iters, P_iters), written after one stream sync.
 * U may alias U_n and P may alias P_n (an in-place step): P is written only after both solves
 * have finished, U only by the final projection. */
int fcfd_four_step_timestep(fcfd_ctx* ctx, const double* U_n, const double* P_n, double dt,
                            double rho, double mu, const double* U_d, const double* F_p,
                            const double* F, double* U, double* P, int32_t* iters);

/* nsteps four_step_timestep calls in one native loop, for a rollout whose compute outputs
 * (F_p, F) and Dirichlet values do not change: the compiled step loop of
 * timestep_loop_scan (simulation.py:362-374, lax.scan over the step) and of timestep_loop
 * (:229-360) when compute_func returns the same (F_p, F) every step.  (U0, P0) -> (U, P) after
 * nsteps steps; the intermediate states alternate between (U, P) and a context buffer pair.
 * iters: host int32[2 * nsteps] (U_iters, P_iters per step) or NULL.  U, P must not alias
 * U0, P0. */
int fcfd_four_step_rollout(fcfd_ctx* ctx, const double* U0, const double* P0, int32_t nsteps,
                           double dt, double rho, double mu, const double* U_d, const double* F_p,
                           const double* F, double* U, double* P, int32_t* iters);

/* ---- passive scalar transport (jax_fem_cfd/solvers/passive_transport.py) ------------ */

/* update_phi (:39-61) or, supg = 1, update_phi_SUPG (:14-37): Crank-Nicolson
 * advection-diffusion of phi from phi_n with Ce(U_n), Ce(U) (beta = 1), diffusion D,
 * Dirichlet values phi_d on the context's gamma_d_phi (device, may be NULL), by BiCGSTAB. */
int fcfd_solve_scalar(fcfd_ctx* ctx, const double* U_n, const double* U, const double* phi_n,
                      double dt, double D, const double* phi_d, int32_t supg, double* phi,
                      int32_t* iters);

/* passive_transport.timestep (:64-97): four_step_timestep then the scalar update.
 * iters: host int32[3] = (U_iters, P_iters, phi_iters). */
int fcfd_passive_timestep(fcfd_ctx* ctx, const double* U_n, const double* P_n, const double* phi_n,
                          double dt, double rho, double mu, double D, const double* U_d,
                          const double* phi_d, const double* F_p, const double* F, int32_t supg,
                          double* U, double* P, double* phi, int32_t* iters);

/* Ctau_tilde_elem_calc (element_calc.py:188-200) of U_n into Ct (ne,4,4) and, when Ce
 * (ne,4,4) is given, calculate_tau (:202-229) into tau (ne).  Either output may be NULL. */
int fcfd_supg_elem_calc(fcfd_ctx* ctx, const double* U_n, const double* Ce, double D, double dt,
                        double r, double* Ct, double* tau);

/* ---- reverse-mode products (adjoint twins of the steps) --------------------------- */

/* Vector-Jacobian product of four_step_timestep (standard.py:148-165) as JAX's reverse-mode
 * AD of the reference defines it: every solve differentiated implicitly through
 * lax.custom_linear_solve (jax_iterative_solvers.py:201-211) -- the momentum solve by BiCGSTAB
 * on the transposed masked operator, the pressure solve by the same MG-PCG (marked symmetric),
 * both from the forward solve's x0 and tol -- and the convection matrix Ce(U_n) differentiated
 * through its (linear) dependence on U_n.  Re-runs the forward step from (U_n, P_n) into U, P
 * (required; must not alias the inputs), then writes gU_n (2nn) and gP_n (nn) for the output
 * cotangents gU (2nn) and gP (nn) (NULL = zero).  The outputs may alias the cotangents.
 * iters (4): forward momentum, forward pressure, adjoint pressure, adjoint momentum.
 * quad4 and quad9 contexts. */
int fcfd_four_step_timestep_vjp(fcfd_ctx* ctx, const double* U_n, const double* P_n, double dt,
                                double rho, double mu, const double* U_d, const double* F_p,
                                const double* F, const double* gU, const double* gP, double* U,
                                double* P, double* gU_n, double* gP_n, int32_t* iters);

/* Vector-Jacobian product of the passive-scalar timestep (passive_transport.py:64-97, plain
 * scalar update or, supg = 1, update_phi_SUPG :14-37): re-runs the step into U, P, phi, then
 * writes gU_n, gP_n, gphi_n for the output cotangents gU, gP, gphi (NULL = zero).  The scalar
 * solve is differentiated like the momentum solve (transposed BiCGSTAB from its forward x0);
 * with SUPG the element matrices are also differentiated through tau, Ctau_tilde(U_n) and the
 * transposed Ce(U_n) mass term (element_calc.py:188-229).
 * iters (6): forward momentum, pressure, scalar; adjoint scalar, pressure, momentum.
 * The reverse sweep of a rollout (timestep_loop_scan, simulation.py:362-374) calls this once
 * per step, last step first, with the states recomputed from checkpoints.  quad4 and quad9
 * contexts, plain and SUPG. */
int fcfd_passive_timestep_vjp(fcfd_ctx* ctx, const double* U_n, const double* P_n,
                              const double* phi_n, double dt, double rho, double mu, double D,
                              const double* U_d, const double* phi_d, const double* F_p,
                              const double* F, int32_t supg, const double* gU, const double* gP,
                              const double* gphi, double* U, double* P, double* phi, double* gU_n,
                              double* gP_n, double* gphi_n, int32_t* iters);

/* ---- BLAS-1 used by the generic cg / bicgstab wrappers ----------------------------- */

/* *out (host) = x . y, FP64, deterministic two-stage reduction */
int fcfd_dot(fcfd_ctx* ctx, const double* x, const double* y, int64_t n, double* out);
/* Largest n accepted by fcfd_dot_noalloc: its block partials (one per 256 elements) live in a
 * fixed device array of the library, 2^16 doubles = 512 KiB per GPU. */
#define FCFD_DOT_NOALLOC_MAX_N (1 << 24)
/* *out (a DEVICE pointer) = x . y with the kernels and summation order of fcfd_dot (bitwise the
 * same value), enqueued on the context's stream; returns without synchronising and allocates
 * nothing: the block partials live in that fixed array, whose address on each GPU is looked up
 * once when the first context on that GPU is created.  n > FCFD_DOT_NOALLOC_MAX_N is rejected
 * with FCFD_ERR_ARG.  Every call on one GPU shares the array: calls that overlap in time on one
 * GPU (other streams, threads or contexts) race.  No counterpart in the reference: this entry
 * point is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
int fcfd_dot_noalloc(fcfd_ctx* ctx, const double* x, const double* y, int64_t n, double* out);
/* y = a x + b y */
int fcfd_axpby(fcfd_ctx* ctx, double a, const double* x, double b, double* y, int64_t n);

/* ---- timing of the dominant kernel (momentum operator apply) ----------------------- */

/* Average device time (ms) of `reps` momentum-operator applications
 * (mass + theta(rho Ce + mu Le), both components, Dirichlet rows) on the
 * context's stored Ce, measured with HIP events on the context stream. */
int fcfd_bench_momentum_apply(fcfd_ctx* ctx, int32_t reps, float* ms_per_apply);

#ifdef __cplusplus
}
#endif
#endif /* FCFD_H */
