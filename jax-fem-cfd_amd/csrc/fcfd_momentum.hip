// Momentum solve: BiCGSTAB (jax_fem_cfd/utils/jax_iterative_solvers.py:123-167) on the fused
// operator (rho/dt) M_c + theta (rho Ce + mu Le) with Dirichlet rows -> identity
// (jax_fem_cfd/solvers/standard.py:14-43), and the projection node kernels of
// four_step_timestep (standard.py:156, :160-163).
//
// The BiCGSTAB kernels are the generic ones of fcfd_bicgstab.hpp with 2 components and the
// fused (cM Me + cL Le + cC Ce_e) element operator.
#include "fcfd_bicgstab.hpp"

namespace fcfd {

// -(G_c P_n) + F_c on the momentum rhs (standard.py:31-33)
struct MomExtra {
    Grid2D g;
    const ElemConst* __restrict__ ec;
    const double *Pn, *F;
    __device__ __forceinline__ void operator()(const NodeNbr& s, int n, double (&b)[2]) const {
        double XP[3][3];
        load3x3(s, [&](int k) { return Pn[k]; }, XP);
        b[0] -= ebe_const(s, ec->Gx, XP);
        b[1] -= ebe_const(s, ec->Gy, XP);
        if (F) { b[0] += F[n]; b[1] += F[n + g.nn]; }
    }
    __device__ __forceinline__ void operator()(const NodeNbr&, int, double (&)[1]) const {}
};

// LDS-staged momentum init (non-periodic grids): BInit<2, MomCoef<S>, MomCoef<S>, MomExtra> with
// the stencils of x_n (both components), x0 (x_n with the Dirichlet values) and P_n staged per
// block window (5 values per node) instead of gathered from global memory: the node-form init
// held its 3x3 stencils of all of them in registers (~160 VGPRs, 3 waves per SIMD).  Operands and
// operation order are BInit's (window entries of wrapped or missing neighbours only reach
// elements that has[k] masks), so every value is bitwise the same.
template <class S>
struct BInitMomS {
    static constexpr int NV = 5;
    static constexpr int NPT = 1;
    MomCoef<S> lhs, rhs;
    const ElemConst* __restrict__ ec;
    const double *Pn, *F, *xn, *ud;
    const uint8_t* mask;
    double *x, *r, *rhat;
    __device__ __forceinline__ int nnx() const { return lhs.g.nnx; }
    __device__ __forceinline__ void stage(int id, double (&v)[5]) const {
        const int nn = lhs.g.nn;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double xv = xn[id + c * nn];
            v[c] = xv;
            v[2 + c] = (mask && mask[id + c * nn]) ? ud[id + c * nn] : xv;
        }
        v[4] = Pn[id];
    }
    __device__ void operator()(int n, const StageView<5, 1>& sv, double (&acc)[2]) const {
        const int nn = lhs.g.nn;
        NodeNbr s;
        node_elems(lhs.g, n, s);
        // stencil values are read from the window where they are used (op_apply_pair's and
        // ebe_const's operation order), not held as 3x3 register arrays
        double b[2] = {0.0, 0.0}, ax[2] = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int a = kRow(k);
            const int e = s.e[k];
            double v[4];
            lhs.ce.ld_row(a, e, v);
            double cl[4], cr[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                cl[q] = lhs.cM * lhs.ec->Me[a * 4 + q] + lhs.cL * lhs.ec->Le[a * 4 + q] + lhs.cC * v[q];
                cr[q] = rhs.cM * rhs.ec->Me[a * 4 + q] + rhs.cL * rhs.ec->Le[a * 4 + q] + rhs.cC * v[q];
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                double tl = 0.0, tr = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int dy = kCy(k, q) - 1, dx = kCx(k, q) - 1;
                    tl += cl[q] * sv.at(2 + c, n, dy, dx);
                    tr += cr[q] * sv.at(c, n, dy, dx);
                }
                ax[c] += s.has[k] ? tl : 0.0;
                b[c] += s.has[k] ? tr : 0.0;
            }
        }
        double gx = 0.0, gy = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int a = kRow(k);
            double tx = 0.0, ty = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double p = sv.at(4, n, kCy(k, q) - 1, kCx(k, q) - 1);
                tx += ec->Gx[a * 4 + q] * p;
                ty += ec->Gy[a * 4 + q] * p;
            }
            gx += s.has[k] ? tx : 0.0;
            gy += s.has[k] ? ty : 0.0;
        }
        b[0] -= gx;
        b[1] -= gy;
        if (F) { b[0] += F[n]; b[1] += F[n + nn]; }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int i = n + c * nn;
            const bool m = mask && mask[i];
            const double bc = m ? ud[i] : b[c];
            const double x0 = sv.at(2 + c, n, 0, 0);
            const double a = m ? x0 : ax[c];
            const double rc = bc - a;
            x[i] = x0;
            r[i] = rc; rhat[i] = rc;
            acc[0] += bc * bc;
            acc[1] += rc * rc;
        }
    }
};

template <class S>
struct MomOps {
    MomCoef<S> lhs, rhs;
    int maxit;
    double key[4];
};
template <class S>
static MomOps<S> mom_ops(Ctx& c, S ce, double rho, double mu, double dt) {
    const double theta = 0.5;
    const Grid2D& g = c.g;
    return MomOps<S>{MomCoef<S>{g, c.ec_dev, ce, rho / dt, theta * mu, theta * rho},
                     MomCoef<S>{g, c.ec_dev, ce, rho / dt, -(1.0 - theta) * mu, -(1.0 - theta) * rho},
                     c.maxit_mom > 0 ? c.maxit_mom : 10 * 2 * g.nn, {rho, mu, dt, 0.0}};
}

// chunk: see bicgstab_iterate (U_hat must then be c.m_x)
template <class S>
static int solve_momentum_t(Ctx& c, S ce, const double* U_n, const double* P_n, double rho, double mu, double dt,
                            const double* F, double* U_hat, int* chunk, ChunkEnd cend) {
    const Grid2D& g = c.g;
    const MomOps<S> o = mom_ops(c, ce, rho, mu, dt);
    int it;
    if (!g.periodic && c.dmask) {
        const dim3 grid((g.nnx + MOMT_C - 1) / MOMT_C, (g.nny + MOMT_R - 1) / MOMT_R);
        const int nb = (int)(grid.x * grid.y);
        if (nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
        k_mom_initt<S, MOMT_C, MOMT_R><<<grid, BLOCK, 0, c.stream>>>(o.lhs, o.rhs, c.ec_dev, P_n, F, U_n, c.ud_full,
                                                                    c.dmask, c.m_x, c.m_r, c.m_rhat, c.partials,
                                                                    c.pstride, c.gate);
        finalize_kernel<2><<<1, FIN_BLOCK, 0, c.stream>>>(nb, c.gate, BInitFin{c.st, c.tol, o.maxit}, c.partials,
                                                          c.pstride);
        FCFD_CHECK_LAUNCH();
        it = bicgstab_iterate<2, false>(c, c.g_mom, o.key, ce.c, o.lhs, c.dmask, o.maxit, c.m_x, c.pred_mom,
                                        chunk, 0, cend);
    } else if (!g.periodic) {
        launch_stencil<2>(c, g.nn, c.gate,
                          BInitMomS<S>{o.lhs, o.rhs, c.ec_dev, P_n, F, U_n, c.ud_full, c.dmask, c.m_x, c.m_r,
                                       c.m_rhat},
                          BInitFin{c.st, c.tol, o.maxit});
        FCFD_CHECK_LAUNCH();
        it = bicgstab_iterate<2, false>(c, c.g_mom, o.key, ce.c, o.lhs, c.dmask, o.maxit, c.m_x, c.pred_mom,
                                        chunk, 0, cend);
    } else {
        it = bicgstab_solve<2>(c, c.g_mom, o.key, ce.c, o.lhs, o.rhs, MomExtra{g, c.ec_dev, P_n, F}, U_n, c.ud_full,
                               c.dmask, o.maxit, c.m_x, c.pred_mom, chunk, cend);
    }
    if (U_hat != c.m_x)
        FCFD_HIP(hipMemcpyAsync(U_hat, c.m_x, sizeof(double) * 2 * g.nn, hipMemcpyDeviceToDevice, c.stream));
    return it;
}

template <class S>
static int resume_momentum_t(Ctx& c, S ce, double rho, double mu, double dt, int launched) {
    const MomOps<S> o = mom_ops(c, ce, rho, mu, dt);
    return bicgstab_iterate<2, false>(c, c.g_mom, o.key, ce.c, o.lhs, c.dmask, o.maxit, c.m_x, c.pred_mom, nullptr,
                                      launched);
}

template <class F>
static int with_ce(Ctx& c, int fmt, void* store, float* scale, F&& f) {
    switch (fmt) {
        case FCFD_CE_F64: return f(CeF64{(double*)store, c.g.ne});
        case FCFD_CE_F32: return f(CeF32{(float*)store, c.g.ne});
        case FCFD_CE_F16S: return f(CeF16S{(__half*)store, scale, c.g.ne});
        case FCFD_CE_BF16S: return f(CeBF16S{(__hip_bfloat16*)store, scale, c.g.ne});
    }
    throw Error(FCFD_ERR_ARG, "unknown Ce format");
}

int solve_momentum(Ctx& c, int fmt, void* store, float* scale, const double* U_n, const double* P_n, double rho,
                   double mu, double dt, const double* F, double* U_hat, int* chunk, ChunkEnd cend) {
    if (chunk && U_hat != c.m_x) throw Error(FCFD_ERR_STATE, "a chunked momentum solve works in place");
    return with_ce(c, fmt, store, scale,
                   [&](auto ce) { return solve_momentum_t(c, ce, U_n, P_n, rho, mu, dt, F, U_hat, chunk, cend); });
}

int resume_momentum(Ctx& c, int fmt, void* store, float* scale, double rho, double mu, double dt, int launched) {
    return with_ce(c, fmt, store, scale, [&](auto ce) { return resume_momentum_t(c, ce, rho, mu, dt, launched); });
}

// ------------------------------------------------------------------------------ benchmark
// The momentum operator apply of BiCGSTAB (K2: s = r - alpha q formed on the fly, t = A s stored,
// 3 dot partials; the tiled kernel on non-periodic grids with Dirichlet nodes, as in the solve),
// run back to back `reps` times under its own kernel symbol (TAG = 1) and timed
// with HIP events on the context stream.  The node kernel alone: its 1-block finalize is not part
// of the timing.
template <class S>
static float bench_t(Ctx& c, S ce, int reps) {
    MomCoef<S> lhs{c.g, c.ec_dev, ce, 2000.0, 0.5e-3, 0.5};
    const int nb = blocks_for(c.g.nn);
    using K = BK2S<2, MomCoef<S>, 1>;
    const int nbs = (c.g.nn + BLOCK * K::NPT - 1) / (BLOCK * K::NPT);
    hipEvent_t e0, e1;
    FCFD_HIP(hipEventCreate(&e0));
    FCFD_HIP(hipEventCreate(&e1));
    const dim3 gt((c.g.nnx + MOMT_C - 1) / MOMT_C, (c.g.nny + MOMT_R - 1) / MOMT_R);
    auto launch = [&]() {
        if (c.g.periodic)
            node_kernel<3><<<nb, BLOCK, 0, c.stream>>>(
                c.g.nn, nullptr, BK2<2, MomCoef<S>, 1>{lhs, c.dmask, c.m_r, c.m_q[0], c.m_t, c.st},
                c.partials, c.pstride);
        else if (c.dmask)
            k_mom_k2t<S, MOMT_C, MOMT_R, 1><<<gt, BLOCK, 0, c.stream>>>(lhs, c.dmask, c.m_r, c.m_q[0], c.m_t, c.st,
                                                                       c.partials, c.pstride, nullptr);
        else
            stencil_kernel<3><<<nbs, BLOCK, 0, c.stream>>>(
                c.g.nn, nullptr, K{lhs, c.dmask, c.m_r, c.m_q[0], c.m_t, c.st}, c.partials, c.pstride);
    };
    launch();
    FCFD_HIP(hipEventRecord(e0, c.stream));
    for (int i = 0; i < reps; ++i) launch();
    FCFD_HIP(hipEventRecord(e1, c.stream));
    FCFD_CHECK_LAUNCH();
    FCFD_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FCFD_HIP(hipEventElapsedTime(&ms, e0, e1));
    FCFD_HIP(hipEventDestroy(e0));
    FCFD_HIP(hipEventDestroy(e1));
    return ms / reps;
}

float bench_momentum_apply(Ctx& c, int reps) {
    switch (c.ce_format) {
        case FCFD_CE_F64: return bench_t(c, CeF64{(double*)c.ce, c.g.ne}, reps);
        case FCFD_CE_F32: return bench_t(c, CeF32{(float*)c.ce, c.g.ne}, reps);
        case FCFD_CE_F16S: return bench_t(c, CeF16S{(__half*)c.ce, c.ce_scale, c.g.ne}, reps);
        case FCFD_CE_BF16S: return bench_t(c, CeBF16S{(__hip_bfloat16*)c.ce, c.ce_scale, c.g.ne}, reps);
    }
    throw Error(FCFD_ERR_ARG, "unknown Ce format");
}

// ------------------------------------------------------------------------------ projection
// U* = U_hat + (dt/rho) G P_n / M   (standard.py:156)
struct UstarBody {
    Grid2D g;
    const ElemConst* ec;
    const double *Uh, *Pn, *M;
    double* Us;
    double c;
    __device__ void operator()(int n, double (&)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3];
        load3x3(s, [&](int k) { return Pn[k]; }, X);
        const double gx = ebe_const(s, ec->Gx, X), gy = ebe_const(s, ec->Gy, X);
        Us[n] = Uh[n] + (c * gx) / M[n];
        Us[n + g.nn] = Uh[n + g.nn] + (c * gy) / M[n + g.nn];
    }
};
void projection_ustar(Ctx& c, const double* U_hat, const double* P_n, double rho, double dt, double* U_star) {
    launch_node<0>(c, c.g.nn, c.gate, UstarBody{c.g, c.ec_dev, U_hat, P_n, c.Mlump, U_star, dt / rho}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

// U = U* - (dt/rho) G P / M ; U[gamma_d] = U_d   (standard.py:160-163)
struct FinalBody {
    Grid2D g;
    const ElemConst* ec;
    const double *Us, *P, *M, *ud;
    const uint8_t* mask;
    double* U;
    double c;
    __device__ void operator()(int n, double (&)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3];
        load3x3(s, [&](int k) { return P[k]; }, X);
        const double gx = ebe_const(s, ec->Gx, X), gy = ebe_const(s, ec->Gy, X);
        double u = Us[n] - (c * gx) / M[n];
        double v = Us[n + g.nn] - (c * gy) / M[n + g.nn];
        if (mask && mask[n]) u = ud[n];
        if (mask && mask[n + g.nn]) v = ud[n + g.nn];
        U[n] = u;
        U[n + g.nn] = v;
    }
};
void projection_final(Ctx& c, const double* U_star, const double* P, double rho, double dt, double* U) {
    launch_node<0>(c, c.g.nn, c.gate, FinalBody{c.g, c.ec_dev, U_star, P, c.Mlump, c.ud_full, c.dmask, U, dt / rho},
                   NoFin{});
    FCFD_CHECK_LAUNCH();
}

}  // namespace fcfd
