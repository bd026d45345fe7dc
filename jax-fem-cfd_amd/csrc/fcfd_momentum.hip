// Momentum solve: BiCGSTAB (jax_fem_cfd/utils/jax_iterative_solvers.py:123-167) on the fused
// operator (rho/dt) M_c + theta (rho Ce + mu Le) with Dirichlet rows -> identity
// (jax_fem_cfd/solvers/standard.py:14-43), and the projection node kernels of
// four_step_timestep (standard.py:156, :160-163).
//
// One BiCGSTAB iteration = 3 kernels:
//   K1  p = r + beta (p - omega q) (formed on the fly at each stencil node, double-buffered),
//       q = A p,  finalize alpha = rho' / (rhat . q)
//   K2  s = r - alpha q (on the fly), t = A s,  finalize exit_early, omega = (t.s)/(t.t)
//   K3  x, r updates,  finalize breakdown codes, next rho', beta, convergence flag
// Scalars live on the device; the host enqueues chunks of iterations and reads the 4-byte
// active flag once per chunk.
#include "fcfd_internal.hpp"
#include "fcfd_ops.hpp"

namespace fcfd {

// ------------------------------------------------------------------------------ init
template <class S>
struct MomInit {
    MomOp<S> lhs, rhs;
    const double *Un, *Pn, *F, *ud;
    const uint8_t* mask;
    double *x, *r, *rhat, *p, *q;
    __device__ void operator()(int n, double (&acc)[2]) const {
        const int nn = lhs.g.nn;
        NodeNbr s;
        node_nbr(lhs.g, n, s);
        double Xu[3][3], Xv[3][3], XP[3][3];
        load3x3(s, [&](int k) { return Un[k]; }, Xu);
        load3x3(s, [&](int k) { return Un[k + nn]; }, Xv);
        load3x3(s, [&](int k) { return Pn[k]; }, XP);
        double bu, bv;
        rhs.apply(s, Xu, Xv, bu, bv);
        bu -= ebe_const(s, lhs.ec->Gx, XP);
        bv -= ebe_const(s, lhs.ec->Gy, XP);
        if (F) { bu += F[n]; bv += F[n + nn]; }
        const bool mu_ = mask && mask[n], mv_ = mask && mask[n + nn];
        if (mu_) bu = ud[n];
        if (mv_) bv = ud[n + nn];
        double X0u[3][3], X0v[3][3];
        if (mask) {
            load3x3(s, [&](int k) { const double a = ud[k], b = Un[k]; return mask[k] ? a : b; }, X0u);
            load3x3(s, [&](int k) { const double a = ud[k + nn], b = Un[k + nn]; return mask[k + nn] ? a : b; }, X0v);
        } else {
            load3x3(s, [&](int k) { return Un[k]; }, X0u);
            load3x3(s, [&](int k) { return Un[k + nn]; }, X0v);
        }
        double au, av;
        lhs.apply(s, X0u, X0v, au, av);
        if (mu_) au = X0u[1][1];
        if (mv_) av = X0v[1][1];
        const double ru = bu - au, rv = bv - av;
        x[n] = X0u[1][1]; x[n + nn] = X0v[1][1];
        r[n] = ru; r[n + nn] = rv;
        rhat[n] = ru; rhat[n + nn] = rv;
        p[n] = ru; p[n + nn] = rv;
        q[n] = ru; q[n + nn] = rv;
        acc[0] += bu * bu + bv * bv;
        acc[1] += ru * ru + rv * rv;
    }
};
struct MomInitFin {
    DevState* st;
    double tol;
    int maxit;
    __device__ void operator()(const double (&t)[2]) const {
        double a2 = tol * tol * t[0];
        st->b_atol2 = a2 > 0.0 ? a2 : 0.0;
        st->b_rho = 1.0; st->b_alpha = 1.0; st->b_omega = 1.0;
        st->b_rho_new = t[1];
        st->b_beta = t[1] / 1.0 * 1.0 / 1.0;
        st->b_k = 0; st->b_ee = 0; st->b_maxit = maxit;
        st->b_active = (t[1] > st->b_atol2) && (0 < maxit);
    }
};

// ------------------------------------------------------------------------------ K1
template <class S>
struct MomK1 {
    MomOp<S> op;
    const uint8_t* mask;
    const double *r, *pold, *qold, *rhat;
    double *pnew, *qnew;
    const DevState* st;
    __device__ void operator()(int n, double (&acc)[1]) const {
        const int nn = op.g.nn;
        const double beta = st->b_beta, omega = st->b_omega;
        NodeNbr s;
        node_nbr(op.g, n, s);
        double Xu[3][3], Xv[3][3];
        load3x3(s, [&](int k) { return r[k] + beta * (pold[k] - omega * qold[k]); }, Xu);
        load3x3(s, [&](int k) { return r[k + nn] + beta * (pold[k + nn] - omega * qold[k + nn]); }, Xv);
        double yu, yv;
        op.apply(s, Xu, Xv, yu, yv);
        if (mask && mask[n]) yu = Xu[1][1];
        if (mask && mask[n + nn]) yv = Xv[1][1];
        pnew[n] = Xu[1][1]; pnew[n + nn] = Xv[1][1];
        qnew[n] = yu; qnew[n + nn] = yv;
        acc[0] += rhat[n] * yu + rhat[n + nn] * yv;
    }
};
struct MomK1Fin {
    DevState* st;
    __device__ void operator()(const double (&t)[1]) const { st->b_alpha = st->b_rho_new / t[0]; }
};

// ------------------------------------------------------------------------------ K2
template <class S>
struct MomK2 {
    MomOp<S> op;
    const uint8_t* mask;
    const double *r, *q;
    double *sv, *tv;
    const DevState* st;
    __device__ void operator()(int n, double (&acc)[3]) const {
        const int nn = op.g.nn;
        const double alpha = st->b_alpha;
        NodeNbr s;
        node_nbr(op.g, n, s);
        double Xu[3][3], Xv[3][3];
        load3x3(s, [&](int k) { return r[k] - alpha * q[k]; }, Xu);
        load3x3(s, [&](int k) { return r[k + nn] - alpha * q[k + nn]; }, Xv);
        double yu, yv;
        op.apply(s, Xu, Xv, yu, yv);
        if (mask && mask[n]) yu = Xu[1][1];
        if (mask && mask[n + nn]) yv = Xv[1][1];
        const double su = Xu[1][1], sv_ = Xv[1][1];
        sv[n] = su; sv[n + nn] = sv_;
        tv[n] = yu; tv[n + nn] = yv;
        acc[0] += su * su + sv_ * sv_;
        acc[1] += yu * su + yv * sv_;
        acc[2] += yu * yu + yv * yv;
    }
};
struct MomK2Fin {
    DevState* st;
    __device__ void operator()(const double (&t)[3]) const {
        st->b_ee = t[0] < st->b_atol2;
        st->b_omega = t[1] / t[2];
    }
};

// ------------------------------------------------------------------------------ K3
struct MomK3 {
    int nn;
    const double *p, *sv, *tv, *rhat;
    double *x, *r;
    const DevState* st;
    __device__ void operator()(int k, double (&acc)[2]) const {
        const double alpha = st->b_alpha, omega = st->b_omega;
        const bool ee = st->b_ee;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int i = k + c * nn;
            const double pi = p[i], si = sv[i], ti = tv[i];
            double xi = x[i];
            double ri;
            if (ee) { xi = xi + alpha * pi; ri = si; }
            else { xi = xi + (alpha * pi + omega * si); ri = si - omega * ti; }
            x[i] = xi; r[i] = ri;
            acc[0] += ri * ri;
            acc[1] += rhat[i] * ri;
        }
    }
};
struct MomK3Fin {
    DevState* st;
    __device__ void operator()(const double (&t)[2]) const {
        const double alpha = st->b_alpha, omega = st->b_omega, rho_cur = st->b_rho_new;
        int k = (omega == 0.0 || alpha == 0.0) ? -11 : st->b_k + 1;
        if (rho_cur == 0.0) k = -10;
        st->b_k = k;
        st->b_rho = rho_cur;
        st->b_rho_new = t[1];
        st->b_beta = t[1] / rho_cur * alpha / omega;
        st->b_active = (t[0] > st->b_atol2) && (k < st->b_maxit) && (k >= 0);
    }
};

// ------------------------------------------------------------------------------ driver
template <class S>
static int solve_momentum_t(Ctx& c, S ce, const double* U_n, const double* P_n, double rho, double mu, double dt,
                            const double* F, double* U_hat) {
    const double theta = 0.5;
    const Grid2D& g = c.g;
    MomOp<S> lhs{g, c.ec_dev, ce, rho / dt, theta * mu, theta * rho};
    MomOp<S> rhs{g, c.ec_dev, ce, rho / dt, -(1.0 - theta) * mu, -(1.0 - theta) * rho};
    const int maxit = c.maxit_mom > 0 ? c.maxit_mom : 10 * 2 * g.nn;
    launch_node<2>(c, g.nn, nullptr,
                   MomInit<S>{lhs, rhs, U_n, P_n, F, c.ud_full, c.dmask, c.m_x, c.m_r, c.m_rhat, c.m_p[0], c.m_q[0]},
                   MomInitFin{c.st, c.tol, maxit});
    FCFD_CHECK_LAUNCH();
    const int* act = &c.st->b_active;
    // one chunk = a pair of iterations, so the double-buffer parity returns to 0
    auto pair = [&]() {
        for (int o = 0; o < 2; ++o) {
            const int nw = o ^ 1;
            launch_node<1>(c, g.nn, act,
                           MomK1<S>{lhs, c.dmask, c.m_r, c.m_p[o], c.m_q[o], c.m_rhat, c.m_p[nw], c.m_q[nw], c.st},
                           MomK1Fin{c.st});
            launch_node<3>(c, g.nn, act, MomK2<S>{lhs, c.dmask, c.m_r, c.m_q[nw], c.m_s, c.m_t, c.st},
                           MomK2Fin{c.st});
            launch_node<2>(c, g.nn, act, MomK3{g.nn, c.m_p[nw], c.m_s, c.m_t, c.m_rhat, c.m_x, c.m_r, c.st},
                           MomK3Fin{c.st});
        }
        FCFD_CHECK_LAUNCH();
    };
    const double key[4] = {rho, mu, dt, 0.0};
    int launched = 0;
    int pairs = c.pred_mom > 0 ? (c.pred_mom + 1) / 2 : 1;
    int h[2] = {0, 0};
    while (true) {
        replay(c, c.g_mom, key, ce.c, pairs, pair);
        launched += 2 * pairs;
        read_state_ints(c, &c.st->b_k, 2, h);  // b_k, b_active
        if (!h[1] || launched >= maxit) break;
        pairs = 1;
    }
    if (U_hat != c.m_x)
        FCFD_HIP(hipMemcpyAsync(U_hat, c.m_x, sizeof(double) * 2 * g.nn, hipMemcpyDeviceToDevice, c.stream));
    if (h[0] > 0) c.pred_mom = h[0];
    return h[0];
}

int solve_momentum(Ctx& c, int fmt, void* store, float* scale, const double* U_n, const double* P_n, double rho,
                   double mu, double dt, const double* F, double* U_hat) {
    switch (fmt) {
        case FCFD_CE_F64:
            return solve_momentum_t(c, CeF64{(double*)store, c.g.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
        case FCFD_CE_F32:
            return solve_momentum_t(c, CeF32{(float*)store, c.g.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
        case FCFD_CE_F16S:
            return solve_momentum_t(c, CeF16S{(__half*)store, scale, c.g.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
        case FCFD_CE_BF16S:
            return solve_momentum_t(c, CeBF16S{(__hip_bfloat16*)store, scale, c.g.ne}, U_n, P_n, rho, mu, dt, F,
                                    U_hat);
    }
    throw Error(FCFD_ERR_ARG, "unknown Ce format");
}

// ------------------------------------------------------------------------------ benchmark
template <class S>
static float bench_t(Ctx& c, S ce, int reps) {
    MomOp<S> lhs{c.g, c.ec_dev, ce, 2000.0, 0.5e-3, 0.5};
    const int nb = blocks_for(c.g.nn);
    hipEvent_t e0, e1;
    FCFD_HIP(hipEventCreate(&e0));
    FCFD_HIP(hipEventCreate(&e1));
    // the node kernel alone (its 1-block finalize is not part of the timed work)
    auto launch = [&]() {
        node_kernel<3><<<nb, BLOCK, 0, c.stream>>>(c.g.nn, nullptr,
                                                   MomK2<S>{lhs, c.dmask, c.m_r, c.m_q[0], c.m_s, c.m_t, c.st},
                                                   c.partials, c.pstride);
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
    launch_node<0>(c, c.g.nn, nullptr, UstarBody{c.g, c.ec_dev, U_hat, P_n, c.Mlump, U_star, dt / rho}, NoFin{});
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
    launch_node<0>(c, c.g.nn, nullptr, FinalBody{c.g, c.ec_dev, U_star, P, c.Mlump, c.ud_full, c.dmask, U, dt / rho}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

}  // namespace fcfd
