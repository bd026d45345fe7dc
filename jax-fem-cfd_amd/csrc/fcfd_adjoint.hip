// Reverse-mode products (vector-Jacobian products) of the Navier-Stokes step
// (jax_fem_cfd/solvers/standard.py:148-165) and of the plain passive-scalar step
// (jax_fem_cfd/solvers/passive_transport.py:64-97), as JAX's reverse-mode AD of the reference's
// rollout (setup/simulation.py:362-374) defines them.  Every linear solve is differentiated
// implicitly, as lax.custom_linear_solve does (utils/jax_iterative_solvers.py:201-211):
//   x = A^-1 b:  bbar = A^-T xbar (same method, tol, preconditioner and x0 as the forward solve;
//                CG is marked symmetric, so its transposed solve is the forward solve),
//                parameter cotangent = -<bbar, dA x>.
// The transposed masked operators are  Ahat^T (mask ? 0 : y) + (mask ? y : 0)  (fcfd_bicgstab.hpp).
//
// Ce(U, beta) is linear in U; the cotangent of U through <w, C(U) x> is gathered per node:
//   g_Id(w, x) = sum_{e at I} sum_q wq [ N_I (sum_a w_a N_a)(sum_b x_b d_d N_b)
//                                        + beta d_d N_I (sum_a w_a N_a)(sum_b x_b N_b) ]
// (d/dU_Id of Ce_A[a][b] = sum_q wq N_a N_I d_d N_b and of Ce_D[a][b] = sum_q wq d_d N_I N_a N_b).
//
// Adjoint of one passive step, with the forward intermediates of the step in the context (U_hat
// in m_x, Ce(U_n, 1/2) in ce, the scalar lhs / rhs element matrices in ade_A / ade_B):
//   1  mu_s = A_s^-T gphi             (BiCGSTAB, transposed stored matrices)      [scalar]
//   2  w = mask_phi ? 0 : mu_s;  gphi_n = B_s^T w;  gU~ = gU - 1/2 g_1(w, phi);
//      gU_n,s = -1/2 g_1(w, phi_n)                                                 [scalar]
//   3  gP~ = gP - dt/rho D((mask ? 0 : gU~) / M), pinned rows zeroed (or its mean recorded)
//   4  mu_p = CG(L, gP~) from the forward x0
//   5  gU* = (mask ? 0 : gU~) + rho/dt G mu_p  (= cotangent of U_hat)
//   6  mu_m = A_m^-T gU*                  (BiCGSTAB, transposed momentum operator)
//   7  w = mask ? 0 : mu_m;  gU_n = R_m^T w - rho sum_c g_1/2(w_c, theta U_hat_c + (1-theta) U_n,c)
//      + gU_n,s;  gP_n = D(dt/rho gU* / M - w)
#include "fcfd_bicgstab.hpp"

namespace fcfd {

// g(w, x) at the node for one component pair (W, X stencils); adds into g[2]
__device__ __forceinline__ void conv_grad_node(const ElemConst* __restrict__ ec, const NodeNbr& s,
                                               const double (&W)[3][3], const double (&X)[3][3], double beta,
                                               double (&g)[2]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int I = kRow(k);
        double gx = 0.0, gy = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double wq = 0.0, xq = 0.0, dx = 0.0, dy = 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double wb = W[kCy(k, b)][kCx(k, b)], xb = X[kCy(k, b)][kCx(k, b)];
                wq += ec->N[q * 4 + b] * wb;
                xq += ec->N[q * 4 + b] * xb;
                dx += ec->dNx[q * 4 + b] * xb;
                dy += ec->dNy[q * 4 + b] * xb;
            }
            const double wwq = ec->wq[q] * wq;
            const double NI = ec->N[q * 4 + I];
            gx += wwq * (NI * dx + beta * ec->dNx[q * 4 + I] * xq);
            gy += wwq * (NI * dy + beta * ec->dNy[q * 4 + I] * xq);
        }
        g[0] += s.has[k] ? gx : 0.0;
        g[1] += s.has[k] ? gy : 0.0;
    }
}

// stencil of a vector with masked entries zeroed
__device__ __forceinline__ void load_masked(const NodeNbr& s, const double* __restrict__ v,
                                            const uint8_t* __restrict__ m, double (&X)[3][3]) {
    if (m)
        load3x3(s, [&](int k) { const double a = v[k]; return m[k] ? 0.0 : a; }, X);
    else
        load3x3(s, [&](int k) { return v[k]; }, X);
}

// ------------------------------------------------------------------------------ step 2 (scalar)
struct AdjScalarBody {
    Grid2D g;
    const ElemConst* __restrict__ ec;
    StoredCoef<CeF64> B;  // rhs element matrices B_e (transposed here)
    const double *mu, *phi, *phin, *gU;
    const uint8_t* mphi;
    double *gphin, *gUt, *gUns;
    const double* gel;  // SUPG element cotangents (k_supg_adj_elem layout) or null
    __device__ void operator()(int n, double (&)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double W[1][3][3], Xp[3][3], Xn[3][3];
        load_masked(s, mu, mphi, W[0]);
        load3x3(s, [&](int k) { return phi[k]; }, Xp);
        load3x3(s, [&](int k) { return phin[k]; }, Xn);
        double y[1];
        op_apply<1>(transpose_op(B), s, W, y);
        gphin[n] = y[0];
        double g1[2] = {0.0, 0.0}, g1n[2] = {0.0, 0.0};
        conv_grad_node(ec, s, W[0], Xp, 1.0, g1);
        conv_grad_node(ec, s, W[0], Xn, 1.0, g1n);
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            double en = 0.0, eu = 0.0;
            if (gel) {
                const size_t ne = g.ne;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const size_t r = (size_t)kRow(k) * 2 + d;
                    const double a = gel[r * ne + s.e[k]], b = gel[(8 + r) * ne + s.e[k]];
                    en += s.has[k] ? a : 0.0;
                    eu += s.has[k] ? b : 0.0;
                }
            }
            const int i = n + d * g.nn;
            gUt[i] = (gU ? gU[i] : 0.0) - 0.5 * g1[d] + eu;
            gUns[i] = -0.5 * g1n[d] + en;
        }
    }
};

// ------------------------------------------------------------------------------ step 3
struct AdjPrhsBody {
    Grid2D g;
    const ElemConst* __restrict__ ec;
    const double *gUt, *M, *gP;
    const uint8_t *dmask, *pin;
    double* out;
    double c;  // dt / rho
    __device__ void operator()(int n, double (&acc)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double Vx[3][3], Vy[3][3];
        const int nn = g.nn;
        if (dmask) {
            load3x3(s, [&](int k) { const double a = gUt[k] / M[k]; return dmask[k] ? 0.0 : a; }, Vx);
            load3x3(s, [&](int k) { const double a = gUt[k + nn] / M[k + nn]; return dmask[k + nn] ? 0.0 : a; }, Vy);
        } else {
            load3x3(s, [&](int k) { return gUt[k] / M[k]; }, Vx);
            load3x3(s, [&](int k) { return gUt[k + nn] / M[k + nn]; }, Vy);
        }
        const double div = ebe_const_T(s, ec->Gx, Vx) + ebe_const_T(s, ec->Gy, Vy);
        double v = (gP ? gP[n] : 0.0) - c * div;
        if (pin && pin[n]) v = 0.0;
        out[n] = v;
        acc[0] += v;
    }
};
struct AdjMeanFin {
    double* slot; int n;
    __device__ void operator()(const double (&t)[1]) const { *slot = t[0] / n; }
};

// ------------------------------------------------------------------------------ step 5
// (the mean removal of mu_p in the pure-Neumann form is dropped: G maps constants to zero)
struct AdjUstarBody {
    Grid2D g;
    const ElemConst* __restrict__ ec;
    const double *gUt, *mup;
    const uint8_t *dmask, *pin;
    const DevState* st;
    double* gUs;
    double c;  // rho / dt
    __device__ void operator()(int n, double (&)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        const bool zero = st->c_zero;
        double X[3][3];
        load3x3(s, [&](int k) { const double a = mup[k]; return (zero || (pin && pin[k])) ? 0.0 : a; }, X);
        const double gx = ebe_const(s, ec->Gx, X), gy = ebe_const(s, ec->Gy, X);
        const int nn = g.nn;
        const double u = (dmask && dmask[n]) ? 0.0 : gUt[n];
        const double v = (dmask && dmask[n + nn]) ? 0.0 : gUt[n + nn];
        gUs[n] = u + c * gx;
        gUs[n + nn] = v + c * gy;
    }
};

// ------------------------------------------------------------------------------ step 7
template <class S>
struct AdjFinalBody {
    Grid2D g;
    const ElemConst* __restrict__ ec;
    MomCoef<S> R;  // rhs operator (rho/dt) M - (1-theta)(rho Ce + mu Le), transposed here
    const double *mum, *Uh, *Un, *gUs, *M, *gUns;
    const uint8_t* dmask;
    double *gUn, *gPn;
    double rho, theta, c;  // c = dt / rho
    __device__ void operator()(int n, double (&)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        const int nn = g.nn;
        double W[2][3][3], X[3][3];
        load_masked(s, mum, dmask, W[0]);
        load_masked(s, mum + nn, dmask ? dmask + nn : nullptr, W[1]);
        double y[2];
        op_apply<2>(transpose_op(R), s, W, y);
        double gc[2] = {0.0, 0.0};
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
            const int o = comp * nn;
            load3x3(s, [&](int k) { return theta * Uh[k + o] + (1.0 - theta) * Un[k + o]; }, X);
            conv_grad_node(ec, s, W[comp], X, 0.5, gc);
        }
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const int i = n + d * nn;
            gUn[i] = y[d] - rho * gc[d] + (gUns ? gUns[i] : 0.0);
        }
        double Zx[3][3], Zy[3][3];
        load3x3(s, [&](int k) { return c * gUs[k] / M[k]; }, Zx);
        load3x3(s, [&](int k) { return c * gUs[k + nn] / M[k + nn]; }, Zy);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                Zx[a][b] -= W[0][a][b];
                Zy[a][b] -= W[1][a][b];
            }
        gPn[n] = ebe_const_T(s, ec->Gx, Zx) + ebe_const_T(s, ec->Gy, Zy);
    }
};

static void ensure_adj(Ctx& c) {
    if (!c.a_gut) {
        const size_t nn = c.g.nn;
        c.a_sx = dalloc<double>(c, nn);
        c.a_gut = dalloc<double>(c, 2 * nn);
        c.a_guns = dalloc<double>(c, 2 * nn);
        c.a_gus = dalloc<double>(c, 2 * nn);
        c.a_mx = dalloc<double>(c, 2 * nn);
    }
}

template <class S>
static int momentum_adjoint(Ctx& c, S ce, const double* U_n, double rho, double mu, double dt, const double* b,
                            double* x) {
    const double theta = 0.5;
    MomCoef<S> lhs{c.g, c.ec_dev, ce, rho / dt, theta * mu, theta * rho};
    const int maxit = c.maxit_mom > 0 ? c.maxit_mom : 10 * 2 * c.g.nn;
    const double key[4] = {rho, mu, dt, 0.0};
    return bicgstab_solve_T<2>(c, c.g_amom, key, ce.c, transpose_op(lhs), b, U_n, c.ud_full, c.dmask, maxit, x,
                               c.pred_mom);
}

template <class S>
static void final_body(Ctx& c, S ce, const double* U_n, double rho, double mu, double dt, const double* gUns,
                       double* gU_n, double* gP_n) {
    const double theta = 0.5;
    MomCoef<S> R{c.g, c.ec_dev, ce, rho / dt, -(1.0 - theta) * mu, -(1.0 - theta) * rho};
    launch_node<0>(c, c.g.nn, nullptr,
                   AdjFinalBody<S>{c.g, c.ec_dev, R, c.a_mx, c.m_x, U_n, c.a_gus, c.Mlump, gUns, c.dmask, gU_n, gP_n,
                                   rho, theta, dt / rho},
                   NoFin{});
    FCFD_CHECK_LAUNCH();
}

template <class F>
static auto with_ce(Ctx& c, F&& f) {
    switch (c.ce_format) {
        case FCFD_CE_F64: return f(CeF64{(double*)c.ce, c.g.ne});
        case FCFD_CE_F32: return f(CeF32{(float*)c.ce, c.g.ne});
        case FCFD_CE_F16S: return f(CeF16S{(__half*)c.ce, c.ce_scale, c.g.ne});
        case FCFD_CE_BF16S: return f(CeBF16S{(__hip_bfloat16*)c.ce, c.ce_scale, c.g.ne});
    }
    throw Error(FCFD_ERR_ARG, "unknown Ce format");
}

// NS step VJP.  Requires the forward step's intermediates in the context (U_hat in m_x,
// Ce(U_n, 1/2) in ce).  gU may already hold the scalar terms (gU~); gUn_extra (or null) is added
// to gU_n.  it[0] = pressure adjoint CG iterations, it[1] = momentum adjoint BiCGSTAB iterations.
void ns_vjp(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* gU,
            const double* gP, const double* gUn_extra, double* gU_n, double* gP_n, int* it) {
    ensure_adj(c);
    const int nn = c.g.nn;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    const double* gUt = gU;
    if (!gUt) {
        launch_fill(c, c.a_gut, 0.0, 2 * nn);
        gUt = c.a_gut;
    }
    // 3: pressure adjoint rhs
    launch_node<1>(c, nn, nullptr, AdjPrhsBody{c.g, c.ec_dev, gUt, c.Mlump, gP, c.dmask, pin, c.c_rhs, dt / rho},
                   AdjMeanFin{&c.st->mean_rhs, nn});
    FCFD_CHECK_LAUNCH();
    // 4: CG on the forward operator from the forward x0 (P_n, pins zeroed)
    const int maxit = c.maxit_p > 0 ? c.maxit_p : 1000;
    bool conv;
    it[0] = pcg(c, c.c_rhs, c.set_zeroP ? nullptr : &c.st->mean_rhs, P_n, c.set_zeroP, c.tol, maxit, &conv, true);
    // 5: cotangent of U_hat
    launch_node<0>(c, nn, nullptr, AdjUstarBody{c.g, c.ec_dev, gUt, c.c_x, c.dmask, pin, c.st, c.a_gus, rho / dt},
                   NoFin{});
    FCFD_CHECK_LAUNCH();
    // 6, 7: momentum adjoint solve and the U_n / P_n cotangents
    with_ce(c, [&](auto ce) {
        it[1] = momentum_adjoint(c, ce, U_n, rho, mu, dt, c.a_gus, c.a_mx);
        final_body(c, ce, U_n, rho, mu, dt, gUn_extra, gU_n, gP_n);
        return 0;
    });
}

// Passive step VJP (plain or SUPG scalar update).  Requires the forward step's intermediates in
// the context (as ns_vjp, plus ade_A / ade_B of this step); U, phi = the step's outputs.
// it[0] = scalar adjoint BiCGSTAB, it[1] = pressure adjoint CG, it[2] = momentum adjoint BiCGSTAB.
void passive_vjp(Ctx& c, const double* U_n, const double* P_n, const double* phi_n, const double* U,
                 const double* phi, double dt, double rho, double mu, double D, int supg, const double* gU,
                 const double* gP, const double* gphi, double* gU_n, double* gP_n, double* gphi_n, int* it) {
    ensure_adj(c);
    const int nn = c.g.nn;
    const double* gUt = gU;
    const double* gUns = nullptr;
    it[0] = 0;
    if (gphi) {
        StoredCoef<CeF64> A{c.g, CeF64{c.ade_A, c.g.ne}};
        const double key[4] = {0.0, 0.0, 0.0, 0.0};
        it[0] = bicgstab_solve_T<1>(c, c.g_aphi, key, c.ade_A, transpose_op(A), gphi, phi_n, c.phi_ud, c.phi_mask,
                                    10 * nn, c.a_sx, c.pred_phi);
        if (supg) {
            if (!c.a_gel) c.a_gel = dalloc<double>(c, 16 * (size_t)c.g.ne);
            supg_adjoint_elem(c, U_n, U, c.a_sx, phi, phi_n, D, dt, c.a_gel);
        }
        launch_node<0>(c, nn, nullptr,
                       AdjScalarBody{c.g, c.ec_dev, StoredCoef<CeF64>{c.g, CeF64{c.ade_B, c.g.ne}}, c.a_sx, phi,
                                     phi_n, gU, c.phi_mask, gphi_n, c.a_gut, c.a_guns, supg ? c.a_gel : nullptr},
                       NoFin{});
        FCFD_CHECK_LAUNCH();
        gUt = c.a_gut;
        gUns = c.a_guns;
    } else {
        launch_fill(c, gphi_n, 0.0, nn);
    }
    int itn[2];
    ns_vjp(c, U_n, P_n, dt, rho, mu, gUt, gP, gUns, gU_n, gP_n, itn);
    it[1] = itn[0];
    it[2] = itn[1];
}

}  // namespace fcfd
