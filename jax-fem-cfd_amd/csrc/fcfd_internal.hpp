// Context, device solver state and host-side launch helpers shared by the
// translation units of libfcfd.
#pragma once
#include <vector>
#include <cstring>
#include <type_traits>
#include <string>
#include <stdexcept>
#include "fcfd_common.hpp"
#include "fcfd_hex.hpp"
#include "../../include/fcfd.h"

namespace fcfd {

constexpr int MAX_LEVELS = 24;
// tiles (nodes per block) of the tiled kernels on non-periodic quad4 grids: the V-cycle's down
// legs (coarse nodes) and up legs, the PCG's K1, the momentum BiCGSTAB's K1 / K2
constexpr int MGT_DOWN_CX = 32, MGT_DOWN_CY = 4, MGT_UP_C = 64, MGT_UP_R = 8;
constexpr int CGT_C = 64, CGT_R = 8, MOMT_C = 64, MOMT_R = 4;
inline int tiles_for(int nx, int ny, int tx, int ty) { return ((nx + tx - 1) / tx) * ((ny + ty - 1) / ty); }
constexpr int MAX_RED = 4;  // reductions per kernel

// Scalars that solver kernels read and the finalize steps of reducing kernels
// write.  Lives in device memory; never read by the host inside a solve except
// for the active flags and iteration counts.
struct DevState {
    // BiCGSTAB (jax_iterative_solvers.py:123-167)
    double b_atol2, b_rho, b_rho_new, b_alpha, b_omega, b_beta;
    int b_k, b_active, b_ee, b_maxit;
    int b_zero, c_zero;  // adjoint solves: the rhs was exactly zero (solution 0)
    // PCG (jax_iterative_solvers.py:87-120)
    double c_atol2, c_gamma, c_alpha, c_beta, c_rr, c_mean_r;
    int c_k, c_active, c_maxit, c_flex;
    // pressure rhs / solution means (pure Neumann form)
    double mean_rhs, mean_p;
    // last coarse-CG iteration count (diagnostic)
    int coarse_k;
    int pad_;
    // generic reduction output
    double dot;
    // speculative steps: go = 0 once a predicted solver chunk of an enqueued step fell short;
    // the step kernels given c.gate then return at entry until the host resumes that step
    int go, go_pad_;
    // multigrid lazy means per level: see MG_* slots
    double mg[MAX_LEVELS][6];
};
enum { MG_R = 0, MG_X = 1, MG_F = 2, MG_C = 3, MG_P = 4, MG_Y = 5 };

// The last finalize of a solver chunk of a speculative step (rec null: any other finalize).
// rec: the step's record in host-coherent memory, {status (0 not run, 1 done, 2 fell short),
// phase that fell short (0 momentum, 1 pressure), momentum iterations, pressure iterations}.
struct ChunkEnd {
    int* rec = nullptr;
    int phase = 0;
};
// Runs after that finalize's own update (active and k are then the solver's flag and count for
// the next iteration): a solve still active has fallen short -- st->go and its flag are cleared,
// so every later gated kernel returns at entry; otherwise its count is recorded.
__device__ inline void chunk_end(DevState* st, const ChunkEnd& ce, int& active, int k) {
    if (active) {
        st->go = 0;
        active = 0;
        ce.rec[1] = ce.phase;
        ce.rec[0] = 2;
        return;
    }
    ce.rec[2 + ce.phase] = k;
    if (ce.phase == 1) ce.rec[0] = 1;
}

// The finalize of a chunk's last iteration is launched ungated (it must record a solve that
// converged earlier in the chunk); its functor checks st->go and the solver flag itself.
template <class Fin>
inline auto fin_gate(const int* active, const Fin& f, int) -> decltype(f.ce.rec, (const int*)nullptr) {
    return f.ce.rec ? nullptr : active;
}
template <class Fin>
inline const int* fin_gate(const int* active, const Fin&, long) {
    return active;
}

struct Level {
    Grid2D g;
    Grid3D g3;                          // 3D contexts: this level's box grid
    ElemConst3* ec3_dev = nullptr;      // 3D contexts: this level's hex8 constants
    ElemConstH27* eh27_dev = nullptr;   // hex27 contexts: this level's 27 x 27 constants
    ElemConst ec;
    ElemConst* ec_dev = nullptr;
    ElemConstQ2* eq2_dev = nullptr;  // quad9 contexts: this level's 9 x 9 constants
    double* Ld = nullptr;       // L_diag
    float* Ld32 = nullptr;
    uint8_t* pin = nullptr;     // pinned form: 1 at gamma_p of this level
    double ldc[3][3] = {{0}};   // L_diag by node class [j class][i class] (first, interior, last)
    std::vector<int64_t> pins;
    // V-cycle vectors, FP64 form
    double *r = nullptr, *x = nullptr, *y = nullptr, *f = nullptr, *y2 = nullptr;
    // FP32 form (mixed precision)
    float *r32 = nullptr, *x32 = nullptr, *y32 = nullptr, *f32 = nullptr, *y32b = nullptr;
};

// One captured chunk of solver iterations (a pair, so the double-buffer parity returns to 0),
// replayed with hipGraphLaunch.  `key` holds the scalars baked into the kernel arguments.
struct IterGraph {
    hipGraphExec_t exec = nullptr;  // a pair of iterations (parity 0 then 1)
    double key[4] = {0, 0, 0, 0};
    const void* kptr = nullptr;
    hipGraphExec_t one[2] = {nullptr, nullptr};  // a single iteration of parity 0 / 1
    double okey[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    const void* okptr[2] = {nullptr, nullptr};
};

struct Ctx {
    hipStream_t stream = nullptr;       // private non-blocking stream every kernel runs on
    hipStream_t user_stream = nullptr;  // the caller's stream (joined with events per call)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipEvent_t ev_spec[2] = {nullptr, nullptr};  // end of the speculative step of each record slot
    // speculative steps: when non-null, the kernels of a step that are not solver iterations take
    // it as their `active` flag (st->go)
    const int* gate = nullptr;
    // set while the iterations of a chunk past the count the host is confident of are enqueued:
    // their kernels with the late gate test take the EARLY instantiations (entry test as well)
    int gate_early = 0;
    // of the iterations of the next speculative chunk, how many the host is confident of (the
    // rest are enqueued with early gates); negative: all of them
    int chunk_sure = -1;
    int hist_mom[3] = {0, 0, 0}, hist_p[3] = {0, 0, 0};  // the last iteration counts
    int spec_chunk_fixed = 0;  // FCFD_SPEC_CHUNK: fixed chunk length of the speculative solves
    IterGraph g_mom, g_cg, g_phi, g_amom, g_aphi;
    bool use_graphs = true;
    bool mg_tiled = true;   // pinned non-periodic quad4 V-cycle: the tiled kernels (L_diag by node class)
    // mixed form, pure Neumann, non-periodic: the FP32 V-cycle runs on the tiled kernels with no
    // pins (zero_pin: nn zeros) and no intermediate mean projections
    bool mg_neu_tiled = false;
    uint8_t* zero_pin = nullptr;
    Grid2D g;
    double lx, ly;
    ElemConst ec;
    ElemConst* ec_dev = nullptr;
    // 3D (hex8) contexts: dim = 3, g3 / ec3 describe the grid; g.nn and g.ne hold its node and
    // element counts so that the dimension-free helpers (reductions, fills, graphs) apply as is
    int dim = 2;
    Grid3D g3{};
    double lz = 0.0;
    ElemConst3 ec3{};
    ElemConst3* ec3_dev = nullptr;
    double* fb_w = nullptr;     // 3D / quad9: per Dirichlet face (edge) its weights sum_q w N_a N_b
    // quad9 (Q2) contexts: order = 2, g is the NODE grid (nnx = 2 nex + 1, or 2 nex periodic)
    int order = 1;
    ElemConstQ2 eq2{};
    ElemConstQ2* eq2_dev = nullptr;
    std::vector<ElemConstH27> eh27;       // hex27 contexts: the 27 x 27 constants (host copy)
    ElemConstH27* eh27_dev = nullptr;
    double tol;
    int set_zeroP, precond, precision, ce_format, maxit_mom, maxit_p;
    int approx_keep = 0;  // approx_inv gamma_p rows: 0 zeroed (standard.py), 1 kept (passive_transport.py)
    int nlev = 0;
    std::vector<Level> lv;
    // Dirichlet data
    int n_gamma_d = 0;
    int64_t* gamma_d_dev = nullptr;
    uint8_t* dmask = nullptr;   // (2nn) 1 on Dirichlet DOFs
    double* ud_full = nullptr;  // (2nn) dense U_d
    double* Mlump = nullptr;    // (2nn) lumped mass
    // pressure pins at the fine level
    uint8_t* ppin = nullptr;    // (nn) or null (Neumann)
    // boundary edges for F_p: node-owner gather lists
    int n_bedge = 0;
    int* fb_off = nullptr;      // (nn+1) CSR offsets node -> incident boundary edge slots
    int* fb_edge = nullptr;     // edge id * 2 + local node
    double* fb_geom = nullptr;  // per edge: nx, ny, length
    int* fb_ui = nullptr;       // per edge slot: index into U_d of u / v at the 2 nodes
    // element matrices (timestep storage)
    void* ce = nullptr;
    float* ce_scale = nullptr;
    double* ce64_api = nullptr;  // FP64 row-blocked buffer for API calls taking (ne,4,4)
    // momentum work (2nn)
    double *m_r = nullptr, *m_rhat = nullptr, *m_p[2] = {nullptr, nullptr}, *m_q[2] = {nullptr, nullptr},
           *m_s = nullptr, *m_t = nullptr, *m_x = nullptr;
    // pressure work (nn)
    double *c_r = nullptr, *c_p[2] = {nullptr, nullptr}, *c_Ap = nullptr, *c_x = nullptr, *c_rhs = nullptr;
    double* c_r2 = nullptr;  // second residual buffer of the PCG whose x / r update runs in the V-cycle
    double* c_z[2] = {nullptr, nullptr};
    double* ustar = nullptr;  // (2nn)
    // coarse CG scratch (coarsest level)
    double *cc_r = nullptr, *cc_p = nullptr, *cc_Ap = nullptr;
    float *cc_r32 = nullptr, *cc_p32 = nullptr, *cc_Ap32 = nullptr;
    float* coarse_inv32 = nullptr;  // mixed form, pinned: dense inverse of the coarse operator
    // reductions
    double* partials = nullptr;
    int pstride = 0;            // partials per reduction slot (max blocks of a reducing kernel)
    double* dot_partials = nullptr;  // the generic dot's own partials (any length; never in a graph)
    int dot_pstride = 0;
    DevState* st = nullptr;
    int* h_pinned = nullptr;   // pinned host mirror for flags / counts
    int *h_spec = nullptr, *d_spec = nullptr;  // speculative-step records (host-coherent): host / device view
    // iteration predictions for chunked launching
    int pred_mom = 4, pred_p = 8, pred_phi = 4;
    // passive scalar: Dirichlet data and the per-element lhs / rhs matrices (lazy)
    int n_gamma_d_phi = 0;
    int64_t* gamma_d_phi_dev = nullptr;
    uint8_t* phi_mask = nullptr;
    double* phi_ud = nullptr;
    double *ade_A = nullptr, *ade_B = nullptr, *s_x = nullptr;
    // adjoint (VJP) work: scalar adjoint solution (nn); U cotangent with the scalar terms, the
    // scalar part of the U_n cotangent, the U_hat cotangent and the momentum adjoint solution (2nn)
    double *a_sx = nullptr, *a_gut = nullptr, *a_guns = nullptr, *a_gus = nullptr, *a_mx = nullptr;
    double *roll_U = nullptr, *roll_P = nullptr;  // fcfd_four_step_rollout's second state
    uint8_t* sweep_mask = nullptr;  // fcfd_jacobi_sweep's pin mask (lazy)
    double* a_gel = nullptr;  // SUPG element cotangents, 16 per element (quad4), 36 (quad9)
    std::vector<void*> allocs;
};

// error handling --------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
#define FCFD_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::fcfd::Error(FCFD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define FCFD_CHECK_LAUNCH() FCFD_HIP(hipGetLastError())
#define FCFD_REQUIRE(cond, msg)                                   \
    do {                                                          \
        if (!(cond)) throw ::fcfd::Error(FCFD_ERR_ARG, msg);      \
    } while (0)

template <class T>
T* dalloc(Ctx& c, size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    FCFD_HIP(hipMalloc(&p, n * sizeof(T)));
    c.allocs.push_back(p);
    return (T*)p;
}

// f(std::true_type{}) for a solver-iteration launch on flag `a` that takes the EARLY instantiation
// (c.gate_early, and a flag), else f(std::false_type{})
template <class F>
inline void with_gate_mode(const Ctx& c, const int* a, F&& f) {
    if (a && c.gate_early)
        f(std::true_type{});
    else
        f(std::false_type{});
}

// sets c.gate_early for its lifetime
struct EarlyGates {
    Ctx& c;
    int old;
    EarlyGates(Ctx& cc, bool on) : c(cc), old(cc.gate_early) { c.gate_early = on ? 1 : 0; }
    ~EarlyGates() { c.gate_early = old; }
};

// Of a speculative chunk of n iterations, how many to enqueue with the late gate only: the first
// c.chunk_sure, or all of them when the host has no count to be sure of or replays captured
// graphs (a captured launch keeps the gate it was captured with).
inline int chunk_sure_of(const Ctx& c, int n) {
    return (c.use_graphs || c.chunk_sure < 0) ? n : (c.chunk_sure < n ? c.chunk_sure : n);
}

// Launch a node kernel over n items and, when it reduces (NR > 0), its finalize kernel.
template <int NR, class Body, class Fin>
inline void launch_node(Ctx& c, int n, const int* active, const Body& body, const Fin& fin) {
    const int nb = blocks_for(n);
    if (NR > 0 && nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
    node_kernel<NR><<<nb, BLOCK, 0, c.stream>>>(n, active, body, c.partials, c.pstride);
    if constexpr (NR > 0)
        finalize_kernel<NR><<<1, FIN_BLOCK, 0, c.stream>>>(nb, fin_gate(active, fin, 0), fin, c.partials, c.pstride);
}

// A non-reducing node kernel in blocks of BS threads: on the small multigrid levels 64-thread
// blocks spread the few nodes over four times as many CUs (each wave's gathers then share a
// CU's load pipeline with fewer waves), which measured faster than 256-thread blocks.
template <int BS, class Body>
inline void launch_node_nored(Ctx& c, int n, const int* active, const Body& body) {
    const int nb = n > 0 ? (n + BS - 1) / BS : 1;
    node_kernel<0, Body, BS><<<nb, BS, 0, c.stream>>>(n, active, body, c.partials, c.pstride);
}

// Same for an LDS-staged stencil kernel (non-periodic grids).
template <int NR, class Body, class Fin>
inline void launch_stencil(Ctx& c, int n, const int* active, const Body& body, const Fin& fin) {
    const int per = BLOCK * Body::NPT;
    const int nb = n > 0 ? (n + per - 1) / per : 1;
    if (NR > 0 && nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
    stencil_kernel<NR><<<nb, BLOCK, 0, c.stream>>>(n, active, body, c.partials, c.pstride);
    if constexpr (NR > 0)
        finalize_kernel<NR><<<1, FIN_BLOCK, 0, c.stream>>>(nb, fin_gate(active, fin, 0), fin, c.partials, c.pstride);
}

// join the private stream to the caller's stream on entry / exit of an API call
void enter(Ctx& c);
void leave(Ctx& c);
void graph_reset(IterGraph& g);

// Run `body` (one chunk of solver iterations, enqueued on c.stream) `reps` times.  With graphs
// enabled the chunk is captured once per key into a HIP graph and replayed, so the host issues
// one launch per chunk instead of one per kernel.
template <class F>
inline void replay(Ctx& c, IterGraph& g, const double (&key)[4], const void* kptr, int reps, F&& body) {
    if (!c.use_graphs) {
        for (int i = 0; i < reps; ++i) body();
        return;
    }
    const bool ok = g.exec && g.kptr == kptr && std::memcmp(g.key, key, sizeof(key)) == 0;
    if (!ok) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);  // the pair graph only (see enqueue_iters)
        g.exec = nullptr;
        hipGraph_t graph = nullptr;
        FCFD_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(c.stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        FCFD_HIP(hipStreamEndCapture(c.stream, &graph));
        hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g.exec = nullptr;
            throw Error(FCFD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        std::memcpy(g.key, key, sizeof(key));
        g.kptr = kptr;
    }
    for (int i = 0; i < reps; ++i) FCFD_HIP(hipGraphLaunch(g.exec, c.stream));
}

// Captures `body` into exec (keyed by key, kptr) on first use, then launches it.
template <class F>
inline void replay_exec(Ctx& c, hipGraphExec_t& exec, double (&gkey)[4], const void*& gkptr, const double (&key)[4],
                        const void* kptr, F&& body) {
    if (!c.use_graphs) {
        body();
        return;
    }
    const bool ok = exec && gkptr == kptr && std::memcmp(gkey, key, sizeof(key)) == 0;
    if (!ok) {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        hipGraph_t graph = nullptr;
        FCFD_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(c.stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        FCFD_HIP(hipStreamEndCapture(c.stream, &graph));
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            exec = nullptr;
            throw Error(FCFD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        std::memcpy(gkey, key, sizeof(key));
        gkptr = kptr;
    }
    FCFD_HIP(hipGraphLaunch(exec, c.stream));
}

// Enqueues n solver iterations, the first of parity p (iteration i uses the double buffers of
// parity (p + i) % 2): a single iteration up to parity 0, then pairs, then a trailing single
// one -- so a chunk is exactly as long as predicted, not rounded up to a pair.
template <class ONE, class PAIR>
inline void enqueue_iters(Ctx& c, IterGraph& g, const double (&key)[4], const void* kptr, int p, int n, ONE&& one,
                          PAIR&& pair) {
    if (n <= 0) return;
    if (p == 1) {
        replay_exec(c, g.one[1], g.okey[1], g.okptr[1], key, kptr, [&]() { one(1); });
        --n;
    }
    if (n >= 2) replay(c, g, key, kptr, n / 2, pair);
    if (n % 2) replay_exec(c, g.one[0], g.okey[0], g.okptr[0], key, kptr, [&]() { one(0); });
}

// host element constants --------------------------------------------------------------------
void quad4_constants(double hx, double hy, ElemConst& ec);

// launch helpers (defined in fcfd_ops.hip) ------------------------------------------------
void launch_fill(Ctx& c, double* x, double v, int n);
void launch_fill32(Ctx& c, float* x, float v, int n);
void build_ce(Ctx& c, const double* U, double beta, int fmt, void* store, float* scale);
void ce_to_api(Ctx& c, const double* entry_major, double* api);
void ce_from_api(Ctx& c, const double* api, double* entry_major);
void elem_mult_const(Ctx& c, int level, const double* A16, const double* x, double* y, int acc);
void build_ce_api(Ctx& c, const double* U, double beta, double* Ce);
void elem_mult_general(Ctx& c, const double* K0_16, const double* A1, double c1, const double* A2, double c2,
                       int ncomp, const double* x, double* y);
void boundary_integral(Ctx& c, const double* U_d, double* F_p);
void mesh_fill(Ctx& c, double* coords, int64_t* nconn);
void scatter_ud(Ctx& c, const double* U_d);
double dot(Ctx& c, const double* x, const double* y, int n);
void dot_to_device(Ctx& c, const double* x, const double* y, int n, double* partials, double* out);
void axpby(Ctx& c, double a, const double* x, double b, double* y, int n);
float bench_momentum_apply(Ctx& c, int reps);

// dense coarse solves (fcfd_capi.hip, fcfd_mg.hip): the inverse of an SPD matrix with
// half-bandwidth w, the pinned rows / columns replaced by the identity, and y = Ainv r (FP32
// entries, FP64 sums; r zeroed at the pins, y 0 there) on an FP64 level
bool spd_inverse_banded(std::vector<double>& A, int n, int w, std::vector<double>& inv, bool symmetrise);
constexpr int COARSE3_DENSE_MAX = 2048;  // 3D coarse levels: at most 16 MB of FP32 inverse
void pin_identity(std::vector<double>& A, int n, const std::vector<int64_t>& pins);
void coarse_dense64(Ctx& c, int nn, const uint8_t* pin, const double* r, double* y, const int* active);

// momentum (fcfd_momentum.hip)
// chunk (non-null, U_hat = c.m_x): only the init and *chunk BiCGSTAB iterations are enqueued
// (*chunk <= 0: the predicted count), no host read; *chunk returns the iterations enqueued and
// resume_momentum completes the solve from there
int solve_momentum(Ctx& c, int fmt, void* store, float* scale, const double* U_n, const double* P_n, double rho,
                   double mu, double dt, const double* F, double* U_hat, int* chunk = nullptr,
                   ChunkEnd ce = ChunkEnd{});
int resume_momentum(Ctx& c, int fmt, void* store, float* scale, double rho, double mu, double dt, int launched);

// passive scalar (fcfd_scalar.hip)
int solve_scalar(Ctx& c, const double* U_n, const double* U, const double* phi_n, double dt, double D,
                 const double* phi_d, int supg, double* phi);
void supg_api(Ctx& c, const double* U_n, const double* Ce_api, double D, double dt, double r, double* Ct, double* tau);

// pressure / multigrid (fcfd_mg.hip)
void vcycle(Ctx& c, const double* r, double* z);
void interpolate_level(Ctx& c, int level, const double* xH, double* xh);
void restrict_level(Ctx& c, int level, const double* xh, double* xH);
void jacobi_sweep(Ctx& c, int level, const double* Le_off16, const double* x, const double* rhs, const double* Ld,
                  const int64_t* gamma_p, int64_t n_gamma_p, int set_zeroP, double* y);
int pressure_cg(Ctx& c, const double* b, double* x, double tol, int maxit, bool* converged);
// The pressure right-hand side c (D U* - F_p), 0 at the pins (standard.py:99-104), for a PCG
// whose init kernel forms it per node instead of reading a stored b
struct PrhsSpec {
    const ElemConst* ec; const double* Us; const double* Fp; double c;
};
// zero_exit: a right-hand side that is exactly zero ends the solve at once with c_zero set
// (the solution is 0; adjoint solves only).  prhs (pinned form only): b is formed by the init
// kernel from prhs (b and bmean are then unused).
// chunk (non-null): only the init and *chunk iterations are enqueued (*chunk <= 0: the predicted
// count), no host read; *chunk returns the iterations enqueued and pcg_resume completes the solve
int pcg(Ctx& c, const double* b, const double* bmean, const double* x0in, int zero_pins, double tol, int maxit,
        bool* converged, bool zero_exit = false, const PrhsSpec* prhs = nullptr, int* chunk = nullptr,
        ChunkEnd ce = ChunkEnd{});
int pcg_resume(Ctx& c, int launched, int maxit, bool* converged);
int solve_pressure(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho,
                   double dt, double* P_out);
// solve_pressure in two parts: the solve (solution in c.c_x) and the output pass into P_out
// chunk: as for pcg; pressure_resume then completes it
// U* = U_hat + (dt/rho) G P_n / M into c.ustar, then pressure_solve_only on it; the pinned form on a
// non-periodic grid forms U* inside the PCG's init kernel (same values)
int ustar_pressure_solve(Ctx& c, const double* U_hat, const double* P_n, const double* F_p, double rho, double dt,
                         int* chunk = nullptr, ChunkEnd ce = ChunkEnd{});
int pressure_solve_only(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                        int* chunk = nullptr, ChunkEnd ce = ChunkEnd{});
int pressure_resume(Ctx& c, int launched);
void pressure_output(Ctx& c, double* P_out);
void projection_ustar(Ctx& c, const double* U_hat, const double* P_n, double rho, double dt, double* U_star);
void projection_final(Ctx& c, const double* U_star, const double* P, double rho, double dt, double* U);
// pressure_output then projection_final from c.ustar (one tiled pass for the pinned form)
void output_and_project(Ctx& c, double* P_out, double rho, double dt, double* U);


// 3D (hex8) path (fcfd_3d.hip)
void hex8_constants(double dx, double dy, double dz, ElemConst3& ec);
void build_ctx3(Ctx& c, const fcfd_mesh_desc* m, const fcfd_solver_desc* s);
void node_sums3(const Grid3D& g, const double (&per_corner)[8], std::vector<double>& out);
// hex27 (triquadratic) path (fcfd_hex27.hip)
void build_ctx_h27(Ctx& c, const fcfd_mesh_desc* m, const fcfd_solver_desc* s);
void four_step_h27(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* U_d,
                   const double* F_p, const double* F, double* U, double* P, int* it);
int solve_momentum_h27(Ctx& c, const double* U_n, const double* P_n, double rho, double mu, double dt, const double* F,
                       double* U_hat);
int solve_pressure_h27(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                       double* P_out);
int pressure_cg_h27(Ctx& c, const double* b, double* x, double tol, int maxit, bool* converged);
void boundary_integral_h27(Ctx& c, const double* U_d, double* F_p);
void build_ce_h27_api(Ctx& c, const double* U, double beta, double* Ce);
void mass_matvec_h27(Ctx& c, const double* x, double* y);
void vcycle_h27_api(Ctx& c, const double* r, double* z);
// 3D multigrid (fcfd_mg3.hip)
void build_mg3(Ctx& c, const fcfd_solver_desc* s);
const double* vcycle3(Ctx& c, const double* r0, const int* active);
void vcycle3_api(Ctx& c, const double* r, double* z);
int pressure_cg3(Ctx& c, const double* b, double* x, double tol, int maxit, bool* converged);
void four_step3(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* U_d,
                const double* F_p, const double* F, double* U, double* P, int* it);
int solve_momentum3(Ctx& c, const double* U_n, const double* P_n, double rho, double mu, double dt, const double* F,
                    double* U_hat);
int solve_pressure3(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                    double* P_out);
void boundary_integral3(Ctx& c, const double* U_d, double* F_p);
void mass_matvec3(Ctx& c, const double* x, double* y);
void build_ce3_api(Ctx& c, const double* U, double beta, double* Ce);
int solve_momentum3_from(Ctx& c, const double* U_n, const double* P_n, double rho, double mu, double dt,
                         const double* F, double* U_hat);

// quad9 (Q2) path (fcfd_q2.hip)
void quad9_constants(double hx, double hy, ElemConstQ2& ec);
void build_ctx_q2(Ctx& c, const fcfd_mesh_desc* m, const fcfd_solver_desc* s);
void four_step_q2(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* U_d,
                  const double* F_p, const double* F, double* U, double* P, int* it);
int solve_momentum_q2(Ctx& c, const double* U_n, const double* P_n, double rho, double mu, double dt, const double* F,
                      double* U_hat);
int solve_pressure_q2(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                      double* P_out);
void boundary_integral_q2(Ctx& c, const double* U_d, double* F_p);
void mass_matvec_q2(Ctx& c, const double* x, double* y);
void build_ce_q2_api(Ctx& c, const double* U, double beta, double* Ce);
void vcycle_q2_api(Ctx& c, const double* r, double* z);
void ns_vjp_q2(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* gU,
               const double* gP, const double* gUn_extra, double* gU_n, double* gP_n, int* it);
void passive_vjp_q2(Ctx& c, const double* U_n, const double* P_n, const double* phi_n, const double* U,
                    const double* phi, double dt, double rho, double mu, double D, int supg, const double* gU,
                    const double* gP, const double* gphi, double* gU_n, double* gP_n, double* gphi_n, int* it);
int solve_scalar_q2(Ctx& c, const double* U_n, const double* U, const double* phi_n, double dt, double D,
                    const double* phi_d, int supg, double* phi);

// reverse-mode products of the steps (fcfd_adjoint.hip)
void ns_vjp(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* gU,
            const double* gP, const double* gUn_extra, double* gU_n, double* gP_n, int* it);
void passive_vjp(Ctx& c, const double* U_n, const double* P_n, const double* phi_n, const double* U,
                 const double* phi, double dt, double rho, double mu, double D, int supg, const double* gU,
                 const double* gP, const double* gphi, double* gU_n, double* gP_n, double* gphi_n, int* it);
void supg_adjoint_elem(Ctx& c, const double* U_n, const double* U, const double* mu, const double* phi,
                       const double* phi_n, double D, double dt, double* gel);

// host sync helper: read device ints into the pinned mirror
void read_state_ints(Ctx& c, const int* dev, int n, int* host);

}  // namespace fcfd
