// Q2 (quad9) four-step timestep on structured 2D grids.
//
// The chain of jax_fem_cfd/solvers/standard.py:148-165 on the meshes of create_mesh_2d_quad9 /
// create_periodic_mesh_2d_quad9 (regular_mesh.py:107-145, :182-218) with the quad9 shape
// functions (shape_functions.py:108-159, 3x3 Gauss) and lin3 boundary edges
// (shape_functions.py:89-106).  The grid of a Q2 context is its NODE grid: Grid2D with
// nnx = 2 nex + 1 (2 nex periodic), nny likewise; element (ie, je) owns the nodes
// (2 ie + a, 2 je + b), a, b in {0, 1, 2}, in the local order
//   0 (0,0)  1 (1,0)  2 (2,0)  3 (2,1)  4 (2,2)  5 (1,2)  6 (0,2)  7 (0,1)  8 (1,1).
//
// Kernels are owner-computes node gathers: node (i, j) has 1, 2 or 4 incident elements (by the
// parity of i and j); it sums, in ascending element order, its row of each, applied to that
// element's 9 nodal values (read directly, so no private array is indexed at run time).
#include <algorithm>
#include <cmath>
#include "fcfd_bicgstab.hpp"
#include "fcfd_cg.hpp"
#include "fcfd_mgnodes.hpp"

namespace fcfd {

// ------------------------------------------------------------------------------ element data
__host__ __device__ constexpr int q9x(int b) { return b == 0 || b == 6 || b == 7 ? 0 : (b == 1 || b == 5 || b == 8 ? 1 : 2); }
__host__ __device__ constexpr int q9y(int b) { return b <= 2 ? 0 : (b == 3 || b == 7 || b == 8 ? 1 : 2); }
__host__ __device__ constexpr int q9loc(int x, int y) {
    return y == 0 ? x : (y == 2 ? 6 - x : (x == 0 ? 7 : (x == 2 ? 3 : 8)));
}

__host__ __device__ __forceinline__ size_t q9_idx(int k, int e, int ne) { return ((size_t)(k / 9) * ne + e) * 9 + k % 9; }

struct Ce9F64 {
    double* c; int ne;
    __device__ __forceinline__ double ld(int k, int e) const { return c[q9_idx(k, e, ne)]; }
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[9]) const {
        const double* p = c + q9_idx(a * 9, e, ne);
#pragma unroll
        for (int b = 0; b < 9; ++b) v[b] = p[b];
    }
    __device__ __forceinline__ void st(int k, int e, double v) const { c[q9_idx(k, e, ne)] = v; }
};
struct Ce9F32 {
    float* c; int ne;
    __device__ __forceinline__ double ld(int k, int e) const { return (double)c[q9_idx(k, e, ne)]; }
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[9]) const {
        const float* p = c + q9_idx(a * 9, e, ne);
#pragma unroll
        for (int b = 0; b < 9; ++b) v[b] = (double)p[b];
    }
    __device__ __forceinline__ void st(int k, int e, double v) const { c[q9_idx(k, e, ne)] = (float)v; }
};
struct Ce9Api {
    double* c;
    __device__ __forceinline__ void st(int k, int e, double v) const { c[(size_t)e * 81 + k] = v; }
};

// quad9 3x3 Gauss data of element 0 and its constants (shape_functions.py:108-159,
// element_calc.py:9-40); q = (eta index) * 3 + (xi index)
void quad9_constants(double hx, double hy, ElemConstQ2& ec) {
    const double g = std::sqrt(3.0 / 5.0);
    const double x1[3] = {-g, 0.0, g}, w1[3] = {5.0 / 9, 8.0 / 9, 5.0 / 9};
    auto L = [](double t, int a) { return a == 0 ? t * (t - 1) / 2 : (a == 1 ? 1 - t * t : t * (t + 1) / 2); };
    auto dL = [](double t, int a) { return a == 0 ? (2 * t - 1) / 2 : (a == 1 ? -2 * t : (2 * t + 1) / 2); };
    for (int qe = 0; qe < 3; ++qe)
        for (int qx = 0; qx < 3; ++qx) {
            const int q = qe * 3 + qx;
            const double xi = x1[qx], eta = x1[qe];
            double dxi[9], deta[9];
            double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
            for (int b = 0; b < 9; ++b) {
                const int a = q9x(b), c = q9y(b);
                ec.N[q * 9 + b] = L(xi, a) * L(eta, c);
                dxi[b] = dL(xi, a) * L(eta, c);
                deta[b] = L(xi, a) * dL(eta, c);
                const double X = a * hx / 2, Y = c * hy / 2;
                J00 += dxi[b] * X; J01 += dxi[b] * Y;
                J10 += deta[b] * X; J11 += deta[b] * Y;
            }
            const double det = J00 * J11 - J01 * J10;
            ec.wq[q] = w1[qx] * w1[qe] * det;
            for (int b = 0; b < 9; ++b) {
                ec.dNx[q * 9 + b] = (J11 * dxi[b] - J01 * deta[b]) / det;
                ec.dNy[q * 9 + b] = (-J10 * dxi[b] + J00 * deta[b]) / det;
            }
        }
    for (int a = 0; a < 9; ++a)
        for (int b = 0; b < 9; ++b) {
            double m = 0, l = 0, gx = 0, gy = 0;
            for (int q = 0; q < 9; ++q) {
                const double w = ec.wq[q], Na = ec.N[q * 9 + a], Nb = ec.N[q * 9 + b];
                m += w * Na * Nb;
                l += w * (ec.dNx[q * 9 + a] * ec.dNx[q * 9 + b] + ec.dNy[q * 9 + a] * ec.dNy[q * 9 + b]);
                gx += w * Na * ec.dNx[q * 9 + b];
                gy += w * Na * ec.dNy[q * 9 + b];
            }
            ec.Me[a * 9 + b] = m;
            ec.Le[a * 9 + b] = l;
            ec.Gx[a * 9 + b] = gx;
            ec.Gy[a * 9 + b] = gy;
        }
    for (int k = 0; k < 81; ++k) ec.Leoff[k] = (k % 10 == 0) ? 0.0 : ec.Le[k];
}

// ------------------------------------------------------------------------------ incidence
// Incident elements of node (i, j): x candidates (i even) element i/2 - 1 at local column 2 and
// element i/2 at column 0, (i odd) element (i-1)/2 at column 1 only; likewise in y.  k = kx + 2 ky
// gives ascending element ids.  Missing elements are clamped to element 0 and flagged.
struct IncQ2 {
    int e[4];
    bool has[4];
    int row[4];
    int ex[4], ey[4];
};

__device__ __forceinline__ void node_inc_q2(const Grid2D& g, int n, IncQ2& s) {
    const int j = n / g.nnx, i = n - j * g.nnx;
    int exs[2], cxs[2], eys[2], cys[2];
    bool vx[2], vy[2];
    const bool ox = i & 1, oy = j & 1;
    exs[0] = ox ? (i - 1) >> 1 : (i >> 1) - 1; cxs[0] = ox ? 1 : 2;
    exs[1] = i >> 1; cxs[1] = 0;
    eys[0] = oy ? (j - 1) >> 1 : (j >> 1) - 1; cys[0] = oy ? 1 : 2;
    eys[1] = j >> 1; cys[1] = 0;
    if (g.periodic) {
        exs[0] = exs[0] < 0 ? exs[0] + g.nex : exs[0];
        eys[0] = eys[0] < 0 ? eys[0] + g.ney : eys[0];
        vx[0] = true; vx[1] = !ox;
        vy[0] = true; vy[1] = !oy;
    } else {
        vx[0] = exs[0] >= 0; vx[1] = !ox && exs[1] < g.nex;
        vy[0] = eys[0] >= 0; vy[1] = !oy && eys[1] < g.ney;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int kx = k & 1, ky = k >> 1;
        const bool h = vx[kx] && vy[ky];
        s.has[k] = h;
        s.ex[k] = h ? exs[kx] : 0;
        s.ey[k] = h ? eys[ky] : 0;
        s.e[k] = s.ey[k] * g.nex + s.ex[k];
        s.row[k] = q9loc(cxs[kx], cys[ky]);
    }
}

// node id of local node b of element (ex, ey)
__device__ __forceinline__ int q9_node(const Grid2D& g, int ex, int ey, int b) {
    int x = 2 * ex + q9x(b), y = 2 * ey + q9y(b);
    if (g.periodic) {
        x = x >= g.nnx ? x - g.nnx : x;
        y = y >= g.nny ? y - g.nny : y;
    }
    return y * g.nnx + x;
}

// sum_k has_k sum_b A(row_k, b, e_k) X(node_b of e_k) for NC components; A from rowf(a, e, cf)
template <int NC, class RowF, class ValF>
__device__ __forceinline__ void q2_apply(const Grid2D& g, const IncQ2& s, RowF rowf, ValF val, double (&y)[NC]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) y[c] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool h = s.has[k];
        const int ex = s.ex[k], ey = s.ey[k], e = s.e[k], a = s.row[k];
        double cf[9];
        rowf(a, e, cf);
        double t[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) t[c] = 0.0;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const int id = q9_node(g, ex, ey, b);
#pragma unroll
            for (int c = 0; c < NC; ++c) t[c] += cf[b] * val(c, id);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] += h ? t[c] : 0.0;
    }
}

// one shared 9 x 9 matrix (TR: its transpose)
template <bool TR = false, class ValF>
__device__ __forceinline__ double q2_const(const Grid2D& g, const IncQ2& s, const double* __restrict__ A, ValF val) {
    double y[1];
    q2_apply<1>(g, s,
                [&](int a, int, double (&cf)[9]) {
#pragma unroll
                    for (int b = 0; b < 9; ++b) cf[b] = TR ? A[b * 9 + a] : A[a * 9 + b];
                },
                [&](int, int id) { return val(id); }, y);
    return y[0];
}

// ------------------------------------------------------------------------------ Ce
// Ce[a][b] = sum_q w N_qa (U_q . grad N_qb + beta div U_q N_qb)   (element_calc.py:165-185)
template <class S>
__global__ __launch_bounds__(BLOCK) void k_build_ce_q2(Grid2D g, const ElemConstQ2* __restrict__ ec,
                                                      const double* __restrict__ U, double beta, S store) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (e >= g.ne) return;
    const int je = e / g.nex, ie = e - je * g.nex;
    double ul[9], vl[9];
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        const int id = q9_node(g, ie, je, b);
        ul[b] = U[id];
        vl[b] = U[id + g.nn];
    }
    double A[81];
#pragma unroll
    for (int k = 0; k < 81; ++k) A[k] = 0.0;
    for (int q = 0; q < 9; ++q) {
        double uq = 0.0, vq = 0.0, div = 0.0;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            uq += ec->N[q * 9 + b] * ul[b];
            vq += ec->N[q * 9 + b] * vl[b];
            div += ec->dNx[q * 9 + b] * ul[b] + ec->dNy[q * 9 + b] * vl[b];
        }
        const double w = ec->wq[q];
        double adv[9];
#pragma unroll
        for (int b = 0; b < 9; ++b) adv[b] = uq * ec->dNx[q * 9 + b] + vq * ec->dNy[q * 9 + b] + beta * div * ec->N[q * 9 + b];
#pragma unroll
        for (int a = 0; a < 9; ++a) {
            const double wNa = w * ec->N[q * 9 + a];
#pragma unroll
            for (int b = 0; b < 9; ++b) A[a * 9 + b] += wNa * adv[b];
        }
    }
#pragma unroll
    for (int k = 0; k < 81; ++k) store.st(k, e, A[k]);
}

static void build_ce_q2(Ctx& c, const double* U, double beta) {
    const int nb = blocks_for(c.g.ne);
    if (c.ce_format == FCFD_CE_F64)
        k_build_ce_q2<<<nb, BLOCK, 0, c.stream>>>(c.g, c.eq2_dev, U, beta, Ce9F64{(double*)c.ce, c.g.ne});
    else
        k_build_ce_q2<<<nb, BLOCK, 0, c.stream>>>(c.g, c.eq2_dev, U, beta, Ce9F32{(float*)c.ce, c.g.ne});
    FCFD_CHECK_LAUNCH();
}

void build_ce_q2_api(Ctx& c, const double* U, double beta, double* Ce) {
    k_build_ce_q2<<<blocks_for(c.g.ne), BLOCK, 0, c.stream>>>(c.g, c.eq2_dev, U, beta, Ce9Api{Ce});
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ momentum
template <class S>
struct MomCoefQ2 {
    Grid2D g;
    const ElemConstQ2* __restrict__ ec;
    S ce;
    double cM, cL, cC;
    __device__ __forceinline__ void row(int a, int e, double (&cf)[9]) const {
        double v[9];
        ce.ld_row(a, e, v);
#pragma unroll
        for (int b = 0; b < 9; ++b) cf[b] = cM * ec->Me[a * 9 + b] + cL * ec->Le[a * 9 + b] + cC * v[b];
    }
};

// -G_c P_n + F_c on the momentum rhs (standard.py:31-33)
struct MomExtraQ2 {
    const ElemConstQ2* ec;
    const double *Pn, *F;
    __device__ __forceinline__ void operator()(const Grid2D& g, const IncQ2& s, int n, double (&b)[2]) const {
        b[0] -= q2_const(g, s, ec->Gx, [&](int id) { return Pn[id]; });
        b[1] -= q2_const(g, s, ec->Gy, [&](int id) { return Pn[id]; });
        if (F) { b[0] += F[n]; b[1] += F[n + g.nn]; }
    }
};
struct NoExtraQ2 {
    template <int NC>
    __device__ __forceinline__ void operator()(const Grid2D&, const IncQ2&, int, double (&)[NC]) const {}
};

// stored per-element 9 x 9 matrices (row-blocked FP64): the scalar-transport operators
struct StoredQ2 {
    Grid2D g;
    Ce9F64 m;
    __device__ __forceinline__ void row(int a, int e, double (&cf)[9]) const { m.ld_row(a, e, cf); }
};

// rhs = R x_n + extra (Dirichlet rows ud), x0 = x_n with Dirichlet values, r0 = rhs - A x0,
// rhat = p = q = r0; reductions b.b, r0.r0 (the quad4 BInit on quad9 gathers)
template <int NC, class LOp, class ROp, class Extra>
struct BInitQ2 {
    LOp lhs;
    ROp rhs;
    Extra extra;
    const double *xn, *ud;
    const uint8_t* mask;
    double *x, *r, *rhat, *p, *q;
    __device__ void operator()(int n, double (&acc)[2]) const {
        const Grid2D& g = lhs.g;
        const int nn = g.nn;
        IncQ2 s;
        node_inc_q2(g, n, s);
        double b[NC], ax[NC];
        q2_apply<NC>(g, s, [&](int a, int e, double (&cf)[9]) { rhs.row(a, e, cf); },
                     [&](int c, int id) { return xn[id + c * nn]; }, b);
        extra(g, s, n, b);
        q2_apply<NC>(g, s, [&](int a, int e, double (&cf)[9]) { lhs.row(a, e, cf); },
                     [&](int c, int id) {
                         const int i = id + c * nn;
                         const double xv = xn[i];
                         return (mask && mask[i]) ? ud[i] : xv;
                     },
                     ax);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const bool m = mask && mask[i];
            const double x0 = m ? ud[i] : xn[i];
            const double bc = m ? ud[i] : b[c];
            const double a = m ? x0 : ax[c];
            const double rc = bc - a;
            x[i] = x0;
            r[i] = rc; rhat[i] = rc; p[i] = rc; q[i] = rc;
            acc[0] += bc * bc;
            acc[1] += rc * rc;
        }
    }
};

// MT: the masked transpose  A^T y = Ahat^T (mask ? 0 : y) + (mask ? y : 0)  of the adjoint solves
template <int NC, class Op, bool MT = false>
struct BK1Q2 {
    Op op;
    const uint8_t* mask;
    const double *r, *pold, *qold, *rhat;
    double *pnew, *qnew;
    const DevState* st;
    __device__ void operator()(int n, double (&acc)[1]) const {
        const Grid2D& g = op.g;
        const int nn = g.nn;
        const double beta = st->b_beta, omega = st->b_omega;
        auto pv = [&](int c, int id) {
            const int i = id + c * nn;
            return r[i] + beta * (pold[i] - omega * qold[i]);
        };
        IncQ2 s;
        node_inc_q2(g, n, s);
        double y[NC];
        q2_apply<NC>(g, s, [&](int a, int e, double (&cf)[9]) { op.row(a, e, cf); },
                     [&](int c, int id) {
                         const double v = pv(c, id);
                         return (MT && mask && mask[id + c * nn]) ? 0.0 : v;
                     },
                     y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const double pc = pv(c, n);
            const bool m = mask && mask[i];
            const double yc = MT ? y[c] + (m ? pc : 0.0) : (m ? pc : y[c]);
            pnew[i] = pc;
            qnew[i] = yc;
            acc[0] += rhat[i] * yc;
        }
    }
};

template <int NC, class Op, bool MT = false>
struct BK2Q2 {
    Op op;
    const uint8_t* mask;
    const double *r, *q;
    double *sv, *tv;
    const DevState* st;
    __device__ void operator()(int n, double (&acc)[3]) const {
        const Grid2D& g = op.g;
        const int nn = g.nn;
        const double alpha = st->b_alpha;
        auto svv = [&](int c, int id) { return r[id + c * nn] - alpha * q[id + c * nn]; };
        IncQ2 s;
        node_inc_q2(g, n, s);
        double y[NC];
        q2_apply<NC>(g, s, [&](int a, int e, double (&cf)[9]) { op.row(a, e, cf); },
                     [&](int c, int id) {
                         const double v = svv(c, id);
                         return (MT && mask && mask[id + c * nn]) ? 0.0 : v;
                     },
                     y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const double sc = svv(c, n);
            const bool m = mask && mask[i];
            const double yc = MT ? y[c] + (m ? sc : 0.0) : (m ? sc : y[c]);
            sv[i] = sc;
            tv[i] = yc;
            acc[0] += sc * sc;
            acc[1] += yc * sc;
            acc[2] += yc * yc;
        }
    }
};

// BiCGSTAB iterations after the init kernel (graph-replayed pairs until the device flag says
// done); returns the reference's iteration count (or -10 / -11)
template <int NC, class LOp, bool MT = false>
static int bicgstab_q2_iterate(Ctx& c, IterGraph& graph, const double (&key)[4], const void* kptr, const LOp& lhs,
                               const uint8_t* mask, int maxit, double* x, int& pred) {
    const int nn = c.g.nn;
    const int* act = &c.st->b_active;
    auto pair = [&]() {
        for (int o = 0; o < 2; ++o) {
            const int nw = o ^ 1;
            launch_node<1>(c, nn, act,
                           BK1Q2<NC, LOp, MT>{lhs, mask, c.m_r, c.m_p[o], c.m_q[o], c.m_rhat, c.m_p[nw], c.m_q[nw],
                                              c.st},
                           BK1Fin{c.st});
            launch_node<3>(c, nn, act, BK2Q2<NC, LOp, MT>{lhs, mask, c.m_r, c.m_q[nw], c.m_s, c.m_t, c.st},
                           BK2Fin{c.st});
            launch_node<2>(c, nn, act, BK3<NC>{nn, c.m_p[nw], c.m_s, c.m_t, c.m_rhat, x, c.m_r, c.st}, BK3Fin{c.st});
        }
        FCFD_CHECK_LAUNCH();
    };
    int launched = 0;
    int pairs = pred > 0 ? (pred + 1) / 2 : 1;
    int h[2] = {0, 0};
    while (true) {
        replay(c, graph, key, kptr, pairs, pair);
        launched += 2 * pairs;
        read_state_ints(c, &c.st->b_k, 2, h);  // b_k, b_active
        if (!h[1] || launched >= maxit) break;
        pairs = 1;
    }
    if (h[0] > 0) pred = h[0];
    return h[0];
}

template <class S>
static int solve_momentum_q2_t(Ctx& c, S ce, const double* U_n, const double* P_n, double rho, double mu, double dt,
                               const double* F, double* U_hat) {
    const double theta = 0.5;
    const Grid2D& g = c.g;
    const int nn = g.nn;
    MomCoefQ2<S> lhs{g, c.eq2_dev, ce, rho / dt, theta * mu, theta * rho};
    MomCoefQ2<S> rhs{g, c.eq2_dev, ce, rho / dt, -(1.0 - theta) * mu, -(1.0 - theta) * rho};
    const int maxit = c.maxit_mom > 0 ? c.maxit_mom : 10 * 2 * nn;
    launch_node<2>(c, nn, nullptr,
                   BInitQ2<2, MomCoefQ2<S>, MomCoefQ2<S>, MomExtraQ2>{lhs, rhs, MomExtraQ2{c.eq2_dev, P_n, F}, U_n,
                                                                     c.ud_full, c.dmask, c.m_x, c.m_r, c.m_rhat,
                                                                     c.m_p[0], c.m_q[0]},
                   BInitFin{c.st, c.tol, maxit});
    FCFD_CHECK_LAUNCH();
    const double key[4] = {rho, mu, dt, 0.0};
    const int it = bicgstab_q2_iterate<2>(c, c.g_mom, key, ce.c, lhs, c.dmask, maxit, c.m_x, c.pred_mom);
    if (U_hat != c.m_x)
        FCFD_HIP(hipMemcpyAsync(U_hat, c.m_x, sizeof(double) * 2 * nn, hipMemcpyDeviceToDevice, c.stream));
    return it;
}

int solve_momentum_q2(Ctx& c, const double* U_n, const double* P_n, double rho, double mu, double dt, const double* F,
                      double* U_hat) {
    build_ce_q2(c, U_n, 0.5);
    if (c.ce_format == FCFD_CE_F64)
        return solve_momentum_q2_t(c, Ce9F64{(double*)c.ce, c.g.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
    return solve_momentum_q2_t(c, Ce9F32{(float*)c.ce, c.g.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
}

// ------------------------------------------------------------------------------ projection
// Uout = Uin + c G P / M (c = dt/rho: U*; c = -dt/rho with Dirichlet values: the final U)
struct ProjQ2Body {
    Grid2D g;
    const ElemConstQ2* ec;
    const double *Uin, *P, *M, *ud;
    const uint8_t* mask;
    double* Uout;
    double c;
    __device__ void operator()(int n, double (&)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        const double gx = q2_const(g, s, ec->Gx, [&](int id) { return P[id]; });
        const double gy = q2_const(g, s, ec->Gy, [&](int id) { return P[id]; });
        const double gd[2] = {gx, gy};
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const int i = n + d * g.nn;
            double u = Uin[i] + (c * gd[d]) / M[i];
            if (mask && mask[i]) u = ud[i];
            Uout[i] = u;
        }
    }
};

// ------------------------------------------------------------------------------ pressure
struct PrhsQ2Body {
    Grid2D g; const ElemConstQ2* ec; const double* Us; const double* Fp; const uint8_t* pin; double* rhs; double c;
    __device__ void operator()(int n, double (&acc)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        const double div = q2_const<true>(g, s, ec->Gx, [&](int id) { return Us[id]; }) +
                           q2_const<true>(g, s, ec->Gy, [&](int id) { return Us[id + g.nn]; });
        double v = c * (div - (Fp ? Fp[n] : 0.0));
        if (pin && pin[n]) v = 0.0;
        rhs[n] = v;
        acc[0] += v;
    }
};

struct CgInitQ2 {
    Grid2D g; const double* Le; const uint8_t* pin; const double* b; const double* bmean; const double* x0in;
    int zero_pins; double* x; double* r;
    __device__ void operator()(int n, double (&acc)[3]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        auto xv = [&](int id) {
            const double v = x0in[id];
            return (zero_pins && pin && pin[id]) ? 0.0 : v;
        };
        const bool pn = pin && pin[n];
        const double ax = pn ? 0.0 : q2_const(g, s, Le, xv);
        const double bn = b[n] - (bmean ? *bmean : 0.0);
        const double rn = bn - ax;
        x[n] = xv(n);
        r[n] = rn;
        acc[0] += bn * bn;
        acc[1] += rn * rn;
        acc[2] += rn;
    }
};

struct CgK1Q2 {
    Grid2D g; const double* Le; const uint8_t* pin; const double* z; const double* pold; double* pnew; double* Ap;
    const DevState* st;
    __device__ void operator()(int n, double (&acc)[1]) const {
        const double beta = st->c_beta;
        auto pv = [&](int id) { return z[id] + beta * pold[id]; };
        IncQ2 s;
        node_inc_q2(g, n, s);
        const double a = (pin && pin[n]) ? 0.0 : q2_const(g, s, Le, pv);
        const double pn = pv(n);
        pnew[n] = pn;
        Ap[n] = a;
        acc[0] += pn * a;
    }
};

// keep: the passive solver's closure, which keeps D1x and D1L at the pins instead of zeroing
// them (passive_transport.py:154-162 against standard.py:225-234)
struct ApproxInvQ2a {
    Grid2D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; double* d1l; int keep;
    __device__ void operator()(int n, double (&)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        const double v = q2_const(g, s, Le, [&](int id) { return r[id] / Ld[id]; }) / Ld[n];
        d1l[n] = (pin && pin[n]) ? (keep ? r[n] / Ld[n] : 0.0) : v;
    }
};
struct ApproxInvQ2b {
    Grid2D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; const double* d1l; double* z;
    int keep;
    __device__ void operator()(int n, double (&)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        double l2 = q2_const(g, s, Le, [&](int id) { return d1l[id]; });
        if (pin && pin[n]) l2 = keep ? d1l[n] : 0.0;
        const double d1x = r[n] / Ld[n];
        z[n] = (3.0 * d1x - 3.0 * d1l[n]) + l2 / Ld[n];
    }
};

// ------------------------------------------------------------------------------ multigrid
// The V-cycle of multigrid.py:231-330 on quad9 levels (FP64): the level operators are the 9 x 9
// quad9 Laplacians, the transfers are the node-grid R = I^T and bilinear prolongation shared
// with quad4 (the quad9 node grid of nex elements is the quad4 node grid of 2 nex).  Neumann
// means are applied lazily through the DevState slots, as in the quad4 V-cycle.

// f = r - L x: pinned (x = 2/3 r / L_diag, 0 at the pins, pinned rows 0; multigrid.py:243-254)
// or Neumann (x, r and f mean-projected; :256-268)
struct DownResQ2 {
    Grid2D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; double* x; double* f;
    const DevState* st; int l; int neumann;
    __device__ void operator()(int n, double (&acc)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        double fv;
        if (neumann) {
            const double mx = st->mg[l][MG_X], mr = st->mg[l][MG_R];
            fv = (r[n] - mr) - q2_const(g, s, Le, [&](int id) { return x[id] - mx; });
        } else {
            auto xv = [&](int id) {
                const double v = (2.0 / 3.0 * r[id]) / Ld[id];
                return (pin && pin[id]) ? 0.0 : v;
            };
            x[n] = xv(n);
            const bool pn = pin && pin[n];
            const double rh = pn ? 0.0 : r[n];
            const double lx = pn ? 0.0 : q2_const(g, s, Le, xv);
            fv = rh - lx;
        }
        f[n] = fv;
        acc[0] += fv;
    }
};

// x_h <- x_h + I x_H, then one damped Jacobi (multigrid.py:214-228, :306-326)
struct JacobiQ2 {
    Grid2D gf, gc; const double* Leoff; const double* Ld; const uint8_t* pin; const double* r; const double* x;
    const double* p; const double* yc; double* y; const DevState* st; int l; int neumann;
    __device__ void operator()(int n, double (&acc)[1]) const {
        IncQ2 s;
        node_inc_q2(gf, n, s);
        double rh, off, xc;
        if (neumann) {
            const double mx = st->mg[l][MG_X], mp = st->mg[l][MG_P];
            auto xv = [&](int id) { return (x[id] - mx) + (p[id] - mp); };
            rh = r[n] - (l == 0 ? 0.0 : st->mg[l][MG_R]);
            off = q2_const(gf, s, Leoff, xv);
            xc = xv(n);
        } else {
            auto xv = [&](int id) {
                const double v = x[id] + interp_at(gf, gc, yc, 0.0, id);
                return (pin && pin[id]) ? 0.0 : v;
            };
            rh = r[n];
            off = q2_const(gf, s, Leoff, xv);
            xc = xv(n);
        }
        double v = (OMEGA * (rh - off)) / Ld[n] + (1.0 - OMEGA) * xc;
        if (!neumann && pin && pin[n]) v = 0.0;
        y[n] = v;
        acc[0] += v;
    }
};

// Pinned coarse level: CG (M = identity) from 0 to tol, maxiter 1000, on L with the pinned rows
// zeroed and the rhs zeroed at the pins (multigrid.py:278-287); one workgroup, vectors in global
// memory (L2-resident), the iteration bounded by maxit.
__global__ __launch_bounds__(1024) void k_coarse_cg_q2(Grid2D g, const double* __restrict__ Le,
                                                      const uint8_t* pin, const double* r, double* x, double* cr,
                                                      double* cp, double* cAp, DevState* st, double tol, int maxit,
                                                      const int* active) {
    if (active && *active == 0) return;
    const int nn = g.nn;
    double g2[1] = {0.0};
    for (int n = threadIdx.x; n < nn; n += blockDim.x) {
        const double b = (pin && pin[n]) ? 0.0 : r[n];
        cr[n] = b; cp[n] = b; x[n] = 0.0;
        g2[0] += b * b;
    }
    block_sum<1>(g2);
    double gamma = g2[0];
    const double atol2 = fmax(tol * tol * gamma, 0.0);
    int k = 0;
    while (gamma > atol2 && k < maxit) {
        double pap[1] = {0.0};
        for (int n = threadIdx.x; n < nn; n += blockDim.x) {
            IncQ2 s;
            node_inc_q2(g, n, s);
            const double a = (pin && pin[n]) ? 0.0 : q2_const(g, s, Le, [&](int id) { return cp[id]; });
            cAp[n] = a;
            pap[0] += cp[n] * a;
        }
        block_sum<1>(pap);
        const double alpha = gamma / pap[0];
        double rr[1] = {0.0};
        for (int n = threadIdx.x; n < nn; n += blockDim.x) {
            x[n] = x[n] + alpha * cp[n];
            const double rn = cr[n] - alpha * cAp[n];
            cr[n] = rn;
            rr[0] += rn * rn;
        }
        block_sum<1>(rr);
        const double beta = rr[0] / gamma;
        for (int n = threadIdx.x; n < nn; n += blockDim.x) cp[n] = cr[n] + beta * cp[n];
        __syncthreads();
        gamma = rr[0];
        ++k;
    }
    for (int n = threadIdx.x; n < nn; n += blockDim.x)
        if (pin && pin[n]) x[n] = 0.0;
    if (threadIdx.x == 0) st->coarse_k = k;
}

static void vcycle_q2(Ctx& c, const double* r0, const int* active) {
    const int N = c.nlev, ld = N - 1;
    const int neu = !c.set_zeroP;
    auto run = [&](int n, const auto& body, const MeanFin& fin) {
        if (neu) launch_node<1>(c, n, active, body, fin);
        else launch_node<0>(c, n, active, body, fin);
    };
    for (int l = 0; l < ld; ++l) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        const int n = L.g.nn;
        const double* rl = l == 0 ? r0 : L.r;
        if (neu)
            launch_node<1>(c, n, active, DownX<double, double>{rl, L.Ld, L.x, c.st, l},
                           MeanFin{&c.st->mg[l][MG_X], nullptr, n});
        run(n, DownResQ2{L.g, L.eq2_dev->Le, rl, L.Ld, L.pin, L.x, L.f, c.st, l, neu},
            MeanFin{&c.st->mg[l][MG_F], nullptr, n});
        run(Lc.g.nn, Restrict<double>{L.g, Lc.g, L.f, Lc.r, c.st, l, neu},
            MeanFin{&c.st->mg[l][MG_C], &c.st->mg[l + 1][MG_R], Lc.g.nn});
    }
    {
        Level& C = c.lv[N - 1];
        const double* rc = N == 1 ? r0 : C.r;
        if (neu)
            k_coarse_jacobi<double, double><<<1, 1024, 0, c.stream>>>(C.g.nn, C.Ld, rc, C.y, c.st, N - 1, active);
        else
            k_coarse_cg_q2<<<1, 1024, 0, c.stream>>>(C.g, C.eq2_dev->Le, C.pin, rc, C.y, c.cc_r, c.cc_p, c.cc_Ap, c.st,
                                                     c.tol, 1000, active);
        FCFD_CHECK_LAUNCH();
    }
    for (int l = ld - 1; l >= 0; --l) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        const int n = L.g.nn;
        const double* rl = l == 0 ? r0 : L.r;
        if (neu)
            launch_node<1>(c, n, active, Prolong<double>{L.g, Lc.g, Lc.y, L.f, c.st, l},
                           MeanFin{&c.st->mg[l][MG_P], nullptr, n});
        run(n, JacobiQ2{L.g, Lc.g, L.eq2_dev->Leoff, L.Ld, L.pin, rl, L.x, L.f, Lc.y, L.y, c.st, l, neu},
            MeanFin{&c.st->mg[l][MG_Y], nullptr, n});
    }
    FCFD_CHECK_LAUNCH();
}

// z = y0 - mean (Neumann) into an API vector
struct ZOutQ2 {
    const double* y; double* z; const DevState* st; int neumann;
    __device__ void operator()(int k, double (&)[1]) const { z[k] = y[k] - (neumann ? st->mg[0][MG_Y] : 0.0); }
};
void vcycle_q2_api(Ctx& c, const double* r, double* z) {
    FCFD_REQUIRE(c.precond == FCFD_PC_MULTIGRID, "context was not created with the multigrid preconditioner");
    const int nn = c.g.nn;
    struct SumBody {
        const double* r;
        __device__ void operator()(int k, double (&acc)[1]) const { acc[0] += r[k]; }
    };
    launch_node<1>(c, nn, nullptr, SumBody{r}, MeanFin{&c.st->mg[0][MG_R], nullptr, nn});
    vcycle_q2(c, r, nullptr);
    launch_node<0>(c, nn, nullptr, ZOutQ2{c.lv[0].y, z, c.st, !c.set_zeroP}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

static const double* precondition_q2(Ctx& c, const int* active) {
    const int nn = c.g.nn;
    switch (c.precond) {
        case FCFD_PC_NONE: return c.c_r;
        case FCFD_PC_JACOBI:
            launch_node<0>(c, nn, active, JacobiPrec{c.c_r, c.lv[0].Ld, c.lv[0].y}, NoFin{});
            return c.lv[0].y;
        case FCFD_PC_APPROX_INV:
            launch_node<0>(c, nn, active,
                           ApproxInvQ2a{c.g, c.eq2_dev->Le, c.c_r, c.lv[0].Ld, c.ppin, c.lv[0].x, c.approx_keep},
                           NoFin{});
            launch_node<0>(c, nn, active,
                           ApproxInvQ2b{c.g, c.eq2_dev->Le, c.c_r, c.lv[0].Ld, c.ppin, c.lv[0].x, c.lv[0].y, c.approx_keep},
                           NoFin{});
            return c.lv[0].y;
        case FCFD_PC_MULTIGRID:
            vcycle_q2(c, c.c_r, active);
            return c.lv[0].y;
    }
    throw Error(FCFD_ERR_UNSUPPORTED, "pressure preconditioner not supported for quad9");
}

// PCG (jax_iterative_solvers.py:87-120); solution in c.c_x
static int pcg_q2(Ctx& c, const double* b, const double* bmean, const double* x0in, int zero_pins, double tol,
                  int maxit, bool* converged, bool zero_exit = false) {
    const Grid2D& g = c.g;
    const int nn = g.nn;
    const int neu = c.precond == FCFD_PC_MULTIGRID && !c.set_zeroP;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    const double* Le = c.eq2_dev->Le;
    launch_node<3>(c, nn, nullptr, CgInitQ2{g, Le, pin, b, bmean, x0in, zero_pins, c.c_x, c.c_r},
                   CgInitFin{c.st, tol, maxit, nn, zero_exit});
    const int* act = &c.st->c_active;
    const double* y0 = precondition_q2(c, act);
    launch_node<1>(c, nn, act, CgGamma0<double>{c.c_r, y0, c.c_z[0], c.c_p[0], c.st, neu}, CgGamma0Fin{c.st});
    FCFD_CHECK_LAUNCH();
    const double* Ld = c.lv[0].Ld;
    auto pair = [&]() {
        for (int o = 0; o < 2; ++o) {
            const int nw = o ^ 1;
            launch_node<1>(c, nn, act, CgK1Q2{g, Le, pin, c.c_z[o], c.c_p[o], c.c_p[nw], c.c_Ap, c.st},
                           CgK1Fin{c.st});
            if (c.precond == FCFD_PC_JACOBI || c.precond == FCFD_PC_NONE) {
                launch_node<2>(c, nn, act,
                               CgK23{c.c_p[nw], c.c_Ap, c.precond == FCFD_PC_JACOBI ? Ld : nullptr, c.c_x, c.c_r,
                                     c.c_z[nw], c.st},
                               CgK23Fin{c.st});
            } else {
                launch_node<2>(c, nn, act, CgK2{c.c_p[nw], c.c_Ap, c.c_x, c.c_r, c.st}, CgK2Fin{c.st, nn});
                const double* y = precondition_q2(c, act);
                launch_node<2>(c, nn, act, CgK3<double>{c.c_r, y, c.c_z[nw], c.c_z[o], c.st, neu, 0},
                               CgK3Fin{c.st, 0});
            }
        }
        FCFD_CHECK_LAUNCH();
    };
    const double key[4] = {0.0, 0.0, 0.0, 0.0};
    int launched = 0;
    int pairs = c.pred_p > 0 ? (c.pred_p + 1) / 2 : 1;
    int h[2] = {0, 0};
    while (true) {
        replay(c, c.g_cg, key, nullptr, pairs, pair);
        launched += 2 * pairs;
        read_state_ints(c, &c.st->c_k, 2, h);  // c_k, c_active
        if (!h[1] || launched >= maxit) break;
        pairs = std::max(1, std::min(16, pairs * 2));
    }
    if (converged) *converged = !h[1] && h[0] <= maxit;
    if (h[0] > 0) c.pred_p = h[0];
    return h[0];
}

int solve_pressure_q2(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                      double* P_out) {
    const int nn = c.g.nn;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    launch_node<1>(c, nn, nullptr, PrhsQ2Body{c.g, c.eq2_dev, U_star, F_p, pin, c.c_rhs, rho / dt},
                   MeanFin{&c.st->mean_rhs, nullptr, nn});
    const int maxit = c.maxit_p > 0 ? c.maxit_p : 1000;
    bool conv;
    const int it = pcg_q2(c, c.c_rhs, c.set_zeroP ? nullptr : &c.st->mean_rhs, P_n, c.set_zeroP, c.tol, maxit, &conv);
    if (!c.set_zeroP) launch_node<1>(c, nn, nullptr, SumXBody{c.c_x}, MeanFin{&c.st->mean_p, nullptr, nn});
    launch_node<0>(c, nn, nullptr, PoutBody{c.c_x, pin, P_out, c.st, !c.set_zeroP}, NoFin{});
    FCFD_CHECK_LAUNCH();
    return it;
}

// ------------------------------------------------------------------------------ mass matvec
struct MassQ2Body {
    Grid2D g; const ElemConstQ2* ec; const double* x; double* y;
    __device__ void operator()(int n, double (&)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double* xc = x + (size_t)d * g.nn;
            y[n + (size_t)d * g.nn] = q2_const(g, s, ec->Me, [&](int id) { return xc[id]; });
        }
    }
};
void mass_matvec_q2(Ctx& c, const double* x, double* y) {
    launch_node<0>(c, c.g.nn, nullptr, MassQ2Body{c.g, c.eq2_dev, x, y}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ boundary integral
// F_p[n] = sum over the node's Dirichlet-edge slots (bconn order) of sum_b W_f[a][b] (n_f . U_d(b)),
// W_f[a][b] = sum_q w_q N_qa N_qb of lin3 on edge f (element_calc.py:111-135)
__global__ void k_boundary_integral_q2(int nn, const int* __restrict__ off, const int* __restrict__ slot,
                                       const double* __restrict__ W, const double* __restrict__ nv,
                                       const int* __restrict__ ui, const double* __restrict__ U_d, double* F_p) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    double acc = 0.0;
    for (int k = off[n]; k < off[n + 1]; ++k) {
        const int f = slot[k] / 3, a = slot[k] - 3 * f;
        double t = 0.0;
        for (int b = 0; b < 3; ++b) {
            const int* u = ui + 6 * (size_t)f + 2 * b;
            t += W[9 * (size_t)f + a * 3 + b] * (nv[2 * f] * U_d[u[0]] + nv[2 * f + 1] * U_d[u[1]]);
        }
        acc += t;
    }
    F_p[n] = acc;
}

void boundary_integral_q2(Ctx& c, const double* U_d, double* F_p) {
    const int nn = c.g.nn;
    if (!c.fb_off) {
        launch_fill(c, F_p, 0.0, nn);
        return;
    }
    k_boundary_integral_q2<<<(nn + 255) / 256, 256, 0, c.stream>>>(nn, c.fb_off, c.fb_edge, c.fb_w, c.fb_geom,
                                                                    c.fb_ui, U_d, F_p);
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ host set-up
template <class T>
static void upload_q2(const std::vector<T>& h, T* d) {
    FCFD_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

// out[n] = sum over the incident elements of per_local[the node's local index], element order
static void node_sums_q2(const Grid2D& g, const double (&per_local)[9], std::vector<double>& out) {
    out.assign(g.nn, 0.0);
    for (int je = 0; je < g.ney; ++je)
        for (int ie = 0; ie < g.nex; ++ie)
            for (int b = 0; b < 9; ++b) {
                int x = 2 * ie + q9x(b), y = 2 * je + q9y(b);
                if (g.periodic) { x %= g.nnx; y %= g.nny; }
                out[y * g.nnx + x] += per_local[b];
            }
}

void build_ctx_q2(Ctx& c, const fcfd_mesh_desc* m, const fcfd_solver_desc* s) {
    FCFD_REQUIRE(m->nex >= 1 && m->ney >= 1, "bad element counts");

    FCFD_REQUIRE(c.ce_format == FCFD_CE_F64 || c.ce_format == FCFD_CE_F32, "quad9 Ce formats: FP64 or FP32");
    c.order = 2;
    Grid2D g;
    g.nex = m->nex; g.ney = m->ney; g.periodic = m->periodic;
    g.nnx = m->periodic ? 2 * m->nex : 2 * m->nex + 1;
    g.nny = m->periodic ? 2 * m->ney : 2 * m->ney + 1;
    g.nn = g.nnx * g.nny;
    g.ne = m->nex * m->ney;
    FCFD_REQUIRE((int64_t)g.nn * 2 < ((int64_t)1 << 31), "grid too large");
    c.g = g;
    c.lx = m->lx; c.ly = m->ly;
    const int nn = g.nn;
    quad9_constants(c.lx / g.nex, c.ly / g.ney, c.eq2);
    c.eq2_dev = dalloc<ElemConstQ2>(c, 1);
    FCFD_HIP(hipMemcpy(c.eq2_dev, &c.eq2, sizeof(ElemConstQ2), hipMemcpyHostToDevice));
    {
        double rs[9];
        for (int a = 0; a < 9; ++a) {
            rs[a] = 0.0;
            for (int b = 0; b < 9; ++b) rs[a] += c.eq2.Me[a * 9 + b];
        }
        std::vector<double> M;
        node_sums_q2(g, rs, M);
        M.resize(2 * (size_t)nn);
        std::copy(M.begin(), M.begin() + nn, M.begin() + nn);
        c.Mlump = dalloc<double>(c, 2 * (size_t)nn);
        upload_q2(M, c.Mlump);
    }
    // levels: quad9 meshes of nex >> l elements (multigrid.py:9-88); level 0 always exists
    const int nlev = c.precond == FCFD_PC_MULTIGRID ? s->mg_levels : 1;
    FCFD_REQUIRE(nlev >= 1 && nlev <= MAX_LEVELS, "bad number of multigrid levels");
    if (nlev > 1)
        FCFD_REQUIRE(g.nex % (1 << (nlev - 1)) == 0 && g.ney % (1 << (nlev - 1)) == 0,
                     "element counts must be divisible by 2^(levels-1)");
    c.nlev = nlev;
    c.lv.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
        Level& L = c.lv[l];
        Grid2D gl;
        gl.nex = g.nex >> l; gl.ney = g.ney >> l; gl.periodic = g.periodic;
        gl.nnx = g.periodic ? 2 * gl.nex : 2 * gl.nex + 1;
        gl.nny = g.periodic ? 2 * gl.ney : 2 * gl.ney + 1;
        gl.nn = gl.nnx * gl.nny;
        gl.ne = gl.nex * gl.ney;
        L.g = gl;
        ElemConstQ2 eq;
        quad9_constants(c.lx / gl.nex, c.ly / gl.ney, eq);
        if (l == 0) eq = c.eq2;
        L.eq2_dev = dalloc<ElemConstQ2>(c, 1);
        FCFD_HIP(hipMemcpy(L.eq2_dev, &eq, sizeof(ElemConstQ2), hipMemcpyHostToDevice));
        double dg[9];
        for (int a = 0; a < 9; ++a) dg[a] = eq.Le[a * 9 + a];
        std::vector<double> Ld;
        node_sums_q2(gl, dg, Ld);
        L.Ld = dalloc<double>(c, gl.nn);
        upload_q2(Ld, L.Ld);
        if (l > 0) L.r = dalloc<double>(c, gl.nn);
        L.x = dalloc<double>(c, gl.nn);
        L.y = dalloc<double>(c, gl.nn);
        L.f = dalloc<double>(c, gl.nn);
        if (c.set_zeroP && c.precond == FCFD_PC_MULTIGRID) {
            std::vector<uint8_t> pin(gl.nn, 0);
            if (s->mg_gamma_p && s->mg_gamma_p_off) {
                for (int64_t k = s->mg_gamma_p_off[l]; k < s->mg_gamma_p_off[l + 1]; ++k) {
                    const int64_t v = s->mg_gamma_p[k];
                    FCFD_REQUIRE(v >= 0 && v < gl.nn, "multigrid pin out of range");
                    pin[v] = 1;
                    L.pins.push_back(v);
                }
            } else {
                pin[0] = 1;
                L.pins.push_back(0);
            }
            L.pin = dalloc<uint8_t>(c, gl.nn);
            upload_q2(pin, L.pin);
        }
    }
    {
        const int ncoarse = c.lv[nlev - 1].g.nn;
        c.cc_r = dalloc<double>(c, ncoarse);
        c.cc_p = dalloc<double>(c, ncoarse);
        c.cc_Ap = dalloc<double>(c, ncoarse);
    }
    if (c.set_zeroP || c.precond == FCFD_PC_APPROX_INV) {
        std::vector<uint8_t> pin(nn, 0);
        if (s->gamma_p && s->n_gamma_p > 0) {
            for (int64_t k = 0; k < s->n_gamma_p; ++k) {
                FCFD_REQUIRE(s->gamma_p[k] >= 0 && s->gamma_p[k] < nn, "gamma_p out of range");
                pin[s->gamma_p[k]] = 1;
            }
        } else {
            pin[0] = 1;
        }
        c.ppin = dalloc<uint8_t>(c, nn);
        upload_q2(pin, c.ppin);
    }
    std::vector<int64_t> gd;
    if (s->gamma_d && s->n_gamma_d > 0) {
        gd.assign(s->gamma_d, s->gamma_d + s->n_gamma_d);
        for (size_t k = 0; k < gd.size(); ++k) {
            FCFD_REQUIRE(gd[k] >= 0 && gd[k] < 2 * (int64_t)nn, "gamma_d out of range");
            FCFD_REQUIRE(k == 0 || gd[k] > gd[k - 1], "gamma_d must be sorted and unique");
        }
        c.n_gamma_d = (int)gd.size();
        std::vector<uint8_t> mask(2 * (size_t)nn, 0);
        for (auto v : gd) mask[v] = 1;
        c.dmask = dalloc<uint8_t>(c, 2 * (size_t)nn);
        upload_q2(mask, c.dmask);
        c.gamma_d_dev = dalloc<int64_t>(c, gd.size());
        upload_q2(gd, c.gamma_d_dev);
        c.ud_full = dalloc<double>(c, 2 * (size_t)nn);
        FCFD_HIP(hipMemset(c.ud_full, 0, sizeof(double) * 2 * nn));
    }
    // passive-scalar Dirichlet nodes
    if (s->gamma_d_phi && s->n_gamma_d_phi > 0) {
        std::vector<int64_t> gp(s->gamma_d_phi, s->gamma_d_phi + s->n_gamma_d_phi);
        std::vector<uint8_t> mask(nn, 0);
        for (auto v : gp) {
            FCFD_REQUIRE(v >= 0 && v < nn, "gamma_d_phi out of range");
            mask[v] = 1;
        }
        c.n_gamma_d_phi = (int)gp.size();
        c.gamma_d_phi_dev = dalloc<int64_t>(c, gp.size());
        upload_q2(gp, c.gamma_d_phi_dev);
        c.phi_mask = dalloc<uint8_t>(c, nn);
        upload_q2(mask, c.phi_mask);
        c.phi_ud = dalloc<double>(c, nn);
        FCFD_HIP(hipMemset(c.phi_ud, 0, sizeof(double) * nn));
    }
    // Dirichlet edges (3 nodes each) -> per-node gather lists in bconn order; lin3 weights from
    // the edge's node coordinates (shape_functions.py:89-106)
    if (s->bconn && s->nvec && s->n_bedge > 0 && !gd.empty()) {
        const int nf = (int)s->n_bedge;
        c.n_bedge = nf;
        std::vector<std::vector<int>> lists(nn);
        std::vector<double> W(9 * (size_t)nf), geom(2 * (size_t)nf);
        std::vector<int> ui(6 * (size_t)nf);
        const double hx = c.lx / g.nex / 2, hy = c.ly / g.ney / 2;
        const double gq = std::sqrt(3.0 / 5.0);
        const double xq[3] = {-gq, 0.0, gq}, wq[3] = {5.0 / 9, 8.0 / 9, 5.0 / 9};
        auto sorted_index = [&](int64_t v) { return (int)(std::lower_bound(gd.begin(), gd.end(), v) - gd.begin()); };
        for (int f = 0; f < nf; ++f) {
            double X[3], Y[3];
            for (int a = 0; a < 3; ++a) {
                const int64_t v = s->bconn[3 * (size_t)f + a];
                FCFD_REQUIRE(v >= 0 && v < nn, "bconn out of range");
                X[a] = (v % g.nnx) * hx;
                Y[a] = (v / g.nnx) * hy;
                for (int d = 0; d < 2; ++d) {
                    const int idx = sorted_index(v + (int64_t)d * nn);
                    FCFD_REQUIRE(idx < (int)gd.size() && gd[idx] == v + (int64_t)d * nn,
                                 "boundary node is not a Dirichlet DOF");
                    ui[6 * (size_t)f + 2 * a + d] = idx;
                }
                lists[v].push_back(f * 3 + a);
            }
            for (int k = 0; k < 9; ++k) W[9 * (size_t)f + k] = 0.0;
            for (int q = 0; q < 3; ++q) {
                const double t = xq[q];
                const double N[3] = {t * (t - 1) / 2, 1 - t * t, t * (t + 1) / 2};
                const double dN[3] = {(2 * t - 1) / 2, -2 * t, (2 * t + 1) / 2};
                const double dx = dN[0] * X[0] + dN[1] * X[1] + dN[2] * X[2];
                const double dy = dN[0] * Y[0] + dN[1] * Y[1] + dN[2] * Y[2];
                const double w = wq[q] * std::sqrt(dx * dx + dy * dy);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) W[9 * (size_t)f + a * 3 + b] += w * N[a] * N[b];
            }
            for (int d = 0; d < 2; ++d) geom[2 * (size_t)f + d] = (double)s->nvec[2 * (size_t)f + d];
        }
        std::vector<int> off(nn + 1, 0), slots;
        for (int n = 0; n < nn; ++n) {
            off[n + 1] = off[n] + (int)lists[n].size();
            slots.insert(slots.end(), lists[n].begin(), lists[n].end());
        }
        c.fb_off = dalloc<int>(c, nn + 1);
        upload_q2(off, c.fb_off);
        c.fb_edge = dalloc<int>(c, slots.size());
        upload_q2(slots, c.fb_edge);
        c.fb_w = dalloc<double>(c, W.size());
        upload_q2(W, c.fb_w);
        c.fb_geom = dalloc<double>(c, geom.size());
        upload_q2(geom, c.fb_geom);
        c.fb_ui = dalloc<int>(c, ui.size());
        upload_q2(ui, c.fb_ui);
    }
    const size_t ne = g.ne;
    if (c.ce_format == FCFD_CE_F64) c.ce = dalloc<double>(c, 81 * ne);
    else c.ce = dalloc<float>(c, 81 * ne);
    const size_t n2 = 2 * (size_t)nn;
    c.m_r = dalloc<double>(c, n2);
    c.m_rhat = dalloc<double>(c, n2);
    for (int k = 0; k < 2; ++k) {
        c.m_p[k] = dalloc<double>(c, n2);
        c.m_q[k] = dalloc<double>(c, n2);
        c.c_p[k] = dalloc<double>(c, nn);
        c.c_z[k] = dalloc<double>(c, nn);
    }
    c.m_s = dalloc<double>(c, n2);
    c.m_t = dalloc<double>(c, n2);
    c.m_x = dalloc<double>(c, n2);
    c.c_r = dalloc<double>(c, nn);
    c.c_Ap = dalloc<double>(c, nn);
    c.c_x = dalloc<double>(c, nn);
    c.c_rhs = dalloc<double>(c, nn);
    c.ustar = dalloc<double>(c, n2);
    c.pstride = blocks_for((int)n2) + 1;
    c.partials = dalloc<double>(c, (size_t)MAX_RED * c.pstride);
    c.st = dalloc<DevState>(c, 1);
    FCFD_HIP(hipMemset(c.st, 0, sizeof(DevState)));
    FCFD_HIP(hipHostMalloc((void**)&c.h_pinned, 64, hipHostMallocDefault));
    FCFD_HIP(hipDeviceSynchronize());
}

// ------------------------------------------------------------------------------ passive scalar
// Per element, from U_n and U (passive_transport.py:64-97, element_calc.py:188-229):
//   Mt = Me + tau Ce(U_n)^T, K = D Le + tau Ct(U_n),
//   A_e = Mt / dt + (Ce(U) + K) / 2,  B_e = Mt / dt - (Ce(U_n) + K) / 2   (Ce with beta = 1)
// tau from calculate_tau with r = 2 (the Frobenius norms of Ce(U) and Ct(U_n), mean |U_n,q|^2 over
// the 9 Gauss points, D + 1e-6); tau = 0 without SUPG.  Row by row, so that no 9 x 9 matrix is
// held in registers: pass 1 forms the norms, pass 2 the rows of A_e and B_e.
__global__ __launch_bounds__(BLOCK) void k_ade_elem_q2(Grid2D g, const ElemConstQ2* __restrict__ ec,
                                                      const double* __restrict__ Un, const double* __restrict__ U,
                                                      double dt, double D, int supg, double* __restrict__ A,
                                                      double* __restrict__ B) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (e >= g.ne) return;
    const int je = e / g.nex, ie = e - je * g.nex;
    double un[9], vn[9], u[9], v[9];
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        const int id = q9_node(g, ie, je, b);
        un[b] = Un[id]; vn[b] = Un[id + g.nn];
        u[b] = U[id]; v[b] = U[id + g.nn];
    }
    // quadrature values: velocities and divergences of U_n and U
    double uq[9], vq[9], dq[9], unq[9], vnq[9], dnq[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const double Nb = ec->N[q * 9 + b], dx = ec->dNx[q * 9 + b], dy = ec->dNy[q * 9 + b];
            a0 += Nb * u[b]; a1 += Nb * v[b]; a2 += dx * u[b] + dy * v[b];
            b0 += Nb * un[b]; b1 += Nb * vn[b]; b2 += dx * un[b] + dy * vn[b];
        }
        uq[q] = a0; vq[q] = a1; dq[q] = a2;
        unq[q] = b0; vnq[q] = b1; dnq[q] = b2;
    }
    // Ce(U)[a][b], Ce(U_n)[a][b] and Ct(U_n)[a][b] (beta = 1)
    auto ce_ab = [&](const double* wu, const double* wv, const double* wd, int a, int b) {
        double t = 0.0;
        for (int q = 0; q < 9; ++q) {
            const double adv = wu[q] * ec->dNx[q * 9 + b] + wv[q] * ec->dNy[q * 9 + b] + wd[q] * ec->N[q * 9 + b];
            t += ec->wq[q] * ec->N[q * 9 + a] * adv;
        }
        return t;
    };
    auto ct_ab = [&](int a, int b) {
        double t = 0.0;
        for (int q = 0; q < 9; ++q) {
            const double ga = unq[q] * ec->dNx[q * 9 + a] + vnq[q] * ec->dNy[q * 9 + a];
            const double gb = unq[q] * ec->dNx[q * 9 + b] + vnq[q] * ec->dNy[q * 9 + b];
            t += ec->wq[q] * ga * gb;
        }
        return t;
    };
    double tau = 0.0;
    if (supg) {
        double nce = 0.0, nct = 0.0, u2 = 0.0;
        for (int a = 0; a < 9; ++a)
            for (int b = 0; b < 9; ++b) {
                const double x = ce_ab(uq, vq, dq, a, b), y = ct_ab(a, b);
                nce += x * x;
                nct += y * y;
            }
        for (int q = 0; q < 9; ++q) u2 += unq[q] * unq[q] + vnq[q] * vnq[q];
        u2 /= 9.0;
        nce = sqrt(nce);
        nct = sqrt(nct);
        const double Dp = D + 1e-6;
        const double Re = (u2 * nce) / (Dp * nct);
        const double t1 = nce / nct, t2 = dt / 2, t3 = t1 * Re;
        tau = pow(1.0 / (t1 * t1) + 1.0 / (t2 * t2) + 1.0 / (t3 * t3), -0.5);
    }
    const size_t ne = g.ne;
    for (int a = 0; a < 9; ++a)
        for (int b = 0; b < 9; ++b) {
            const int k = a * 9 + b;
            const double cu = ce_ab(uq, vq, dq, a, b), cn = ce_ab(unq, vnq, dnq, a, b);
            double Mt = ec->Me[k], K = D * ec->Le[k];
            if (supg) {
                Mt = Mt + tau * ce_ab(unq, vnq, dnq, b, a);
                K = K + tau * ct_ab(a, b);
            }
            A[q9_idx(k, e, (int)ne)] = (1.0 / dt) * Mt + 0.5 * (cu + K);
            B[q9_idx(k, e, (int)ne)] = (1.0 / dt) * Mt - 0.5 * (cn + K);
        }
}

__global__ void k_scatter_phi_q2(int n, const int64_t* __restrict__ gd, const double* __restrict__ v, double* full) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) full[gd[k]] = v ? v[k] : 0.0;
}

int solve_scalar_q2(Ctx& c, const double* U_n, const double* U, const double* phi_n, double dt, double D,
                    const double* phi_d, int supg, double* phi) {
    const size_t ne = c.g.ne;
    const int nn = c.g.nn;
    if (!c.ade_A) {
        c.ade_A = dalloc<double>(c, 81 * ne);
        c.ade_B = dalloc<double>(c, 81 * ne);
        c.s_x = dalloc<double>(c, nn);
    }
    k_ade_elem_q2<<<blocks_for(c.g.ne), BLOCK, 0, c.stream>>>(c.g, c.eq2_dev, U_n, U, dt, D, supg, c.ade_A, c.ade_B);
    FCFD_CHECK_LAUNCH();
    if (c.n_gamma_d_phi > 0) {
        k_scatter_phi_q2<<<(c.n_gamma_d_phi + 255) / 256, 256, 0, c.stream>>>(c.n_gamma_d_phi, c.gamma_d_phi_dev,
                                                                               phi_d, c.phi_ud);
        FCFD_CHECK_LAUNCH();
    }
    StoredQ2 lhs{c.g, Ce9F64{c.ade_A, c.g.ne}}, rhs{c.g, Ce9F64{c.ade_B, c.g.ne}};
    const int maxit = 10 * nn;
    launch_node<2>(c, nn, nullptr,
                   BInitQ2<1, StoredQ2, StoredQ2, NoExtraQ2>{lhs, rhs, NoExtraQ2{}, phi_n, c.phi_ud, c.phi_mask, c.s_x,
                                                            c.m_r, c.m_rhat, c.m_p[0], c.m_q[0]},
                   BInitFin{c.st, c.tol, maxit});
    FCFD_CHECK_LAUNCH();
    const double key[4] = {0.0, 0.0, 0.0, 0.0};
    const int it = bicgstab_q2_iterate<1>(c, c.g_phi, key, c.ade_A, lhs, c.phi_mask, maxit, c.s_x, c.pred_phi);
    if (phi != c.s_x)
        FCFD_HIP(hipMemcpyAsync(phi, c.s_x, sizeof(double) * nn, hipMemcpyDeviceToDevice, c.stream));
    return it;
}

// ------------------------------------------------------------------------------ reverse mode (NS step)
// The NS-step VJP of the quad4 path (fcfd_adjoint.hip: every solve differentiated implicitly as
// lax.custom_linear_solve does, utils/jax_iterative_solvers.py:201-211) on quad9 gathers:
//   3  gP~ = gP - dt/rho D((mask ? 0 : gU) / M), pinned rows zeroed (or its mean recorded)
//   4  mu_p = CG(L, gP~) from the forward x0 (CG is symmetric: the transposed solve is the forward)
//   5  gU* = (mask ? 0 : gU) + rho/dt G mu_p
//   6  mu_m = A_m^-T gU*  (BiCGSTAB on the masked transpose, from the forward x0)
//   7  w = mask ? 0 : mu_m;  gU_n = R_m^T w - rho sum_c g_1/2(w_c, theta U_hat_c + (1-theta) U_n,c);
//      gP_n = D(dt/rho gU* / M - w)
// g_beta(w, x) is the cotangent of U through <w, Ce(U, beta) x> (Ce is linear in U), gathered per
// node over its incident elements at the 9 Gauss points.

// transposed rows: row a of the transpose = column a of (cM Me + cL Le + cC Ce_e)
template <class S>
struct MomCoefQ2T {
    Grid2D g;
    const ElemConstQ2* __restrict__ ec;
    S ce;
    double cM, cL, cC;
    __device__ __forceinline__ void row(int a, int e, double (&cf)[9]) const {
#pragma unroll
        for (int b = 0; b < 9; ++b)
            cf[b] = cM * ec->Me[b * 9 + a] + cL * ec->Le[b * 9 + a] + cC * ce.ld(b * 9 + a, e);
    }
};

// x0 = x_n with Dirichlet values, r0 = b - A^T x0 (masked transpose), rhat = p = q = r0;
// reductions b.b, r0.r0
template <int NC, class LOp>
struct BInitVecQ2 {
    LOp lhs;
    const double *bvec, *xn, *ud;
    const uint8_t* mask;
    double *x, *r, *rhat, *p, *q;
    __device__ void operator()(int n, double (&acc)[2]) const {
        const Grid2D& g = lhs.g;
        const int nn = g.nn;
        IncQ2 s;
        node_inc_q2(g, n, s);
        auto x0 = [&](int c, int id) {
            const int i = id + c * nn;
            const double xv = xn[i];
            return (mask && mask[i]) ? ud[i] : xv;
        };
        double ax[NC];
        q2_apply<NC>(g, s, [&](int a, int e, double (&cf)[9]) { lhs.row(a, e, cf); },
                     [&](int c, int id) {
                         const double v = x0(c, id);
                         return (mask && mask[id + c * nn]) ? 0.0 : v;
                     },
                     ax);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const double xc = x0(c, n);
            const double a = ax[c] + ((mask && mask[i]) ? xc : 0.0);
            const double bc = bvec[i];
            const double rc = bc - a;
            x[i] = xc;
            r[i] = rc; rhat[i] = rc; p[i] = rc; q[i] = rc;
            acc[0] += bc * bc;
            acc[1] += rc * rc;
        }
    }
};

// candidate k (kx = k & 1, ky = k >> 1) of node (i, j), formed arithmetically (node_inc_q2's rules)
__device__ __forceinline__ void q2_cand(const Grid2D& g, int i, int j, int k, int& ex, int& ey, int& row, bool& has) {
    const int kx = k & 1, ky = k >> 1;
    const bool ox = i & 1, oy = j & 1;
    ex = kx ? (i >> 1) : (ox ? (i - 1) >> 1 : (i >> 1) - 1);
    ey = ky ? (j >> 1) : (oy ? (j - 1) >> 1 : (j >> 1) - 1);
    const int cx = kx ? 0 : (ox ? 1 : 2), cy = ky ? 0 : (oy ? 1 : 2);
    bool vx, vy;
    if (g.periodic) {
        ex = ex < 0 ? ex + g.nex : ex;
        ey = ey < 0 ? ey + g.ney : ey;
        vx = kx ? !ox : true;
        vy = ky ? !oy : true;
    } else {
        vx = kx ? (!ox && ex < g.nex) : ex >= 0;
        vy = ky ? (!oy && ey < g.ney) : ey >= 0;
    }
    has = vx && vy;
    row = q9loc(cx, cy);
}

// g_beta(w, x) at node n for one component pair: adds into g[2].  The candidate loop is rolled
// (candidates formed arithmetically): one element's 9 + 9 values are live at a time.
template <class WF, class XF>
__device__ __forceinline__ void conv_grad_q2(const Grid2D& g, const ElemConstQ2* __restrict__ ec, int n, WF w, XF x,
                                             double beta, double (&gd)[2]) {
    const int j = n / g.nnx, i = n - j * g.nnx;
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        int ex, ey, I;
        bool has;
        q2_cand(g, i, j, k, ex, ey, I, has);
        if (!has) continue;
        double wv[9], xv[9];
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const int id = q9_node(g, ex, ey, b);
            wv[b] = w(id);
            xv[b] = x(id);
        }
        double gx = 0.0, gy = 0.0;
        for (int q = 0; q < 9; ++q) {
            double wq = 0.0, xq = 0.0, dx = 0.0, dy = 0.0;
#pragma unroll
            for (int b = 0; b < 9; ++b) {
                wq += ec->N[q * 9 + b] * wv[b];
                xq += ec->N[q * 9 + b] * xv[b];
                dx += ec->dNx[q * 9 + b] * xv[b];
                dy += ec->dNy[q * 9 + b] * xv[b];
            }
            const double wwq = ec->wq[q] * wq;
            const double NI = ec->N[q * 9 + I];
            gx += wwq * (NI * dx + beta * ec->dNx[q * 9 + I] * xq);
            gy += wwq * (NI * dy + beta * ec->dNy[q * 9 + I] * xq);
        }
        gd[0] += gx;
        gd[1] += gy;
    }
}

struct AdjPrhsQ2 {
    Grid2D g;
    const ElemConstQ2* __restrict__ ec;
    const double *gUt, *M, *gP;
    const uint8_t *dmask, *pin;
    double* out;
    double c;  // dt / rho
    __device__ void operator()(int n, double (&acc)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        const int nn = g.nn;
        auto v = [&](int d, int id) {
            const int i = id + d * nn;
            const double a = gUt[i] / M[i];
            return (dmask && dmask[i]) ? 0.0 : a;
        };
        const double div = q2_const<true>(g, s, ec->Gx, [&](int id) { return v(0, id); }) +
                           q2_const<true>(g, s, ec->Gy, [&](int id) { return v(1, id); });
        double r = (gP ? gP[n] : 0.0) - c * div;
        if (pin && pin[n]) r = 0.0;
        out[n] = r;
        acc[0] += r;
    }
};

// (the mean removal of mu_p in the pure-Neumann form is dropped: G maps constants to zero)
struct AdjUstarQ2 {
    Grid2D g;
    const ElemConstQ2* __restrict__ ec;
    const double *gUt, *mup;
    const uint8_t *dmask, *pin;
    const DevState* st;
    double* gUs;
    double c;  // rho / dt
    __device__ void operator()(int n, double (&)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        const bool zero = st->c_zero;
        auto mv = [&](int id) {
            const double a = mup[id];
            return (zero || (pin && pin[id])) ? 0.0 : a;
        };
        const double gx = q2_const(g, s, ec->Gx, mv), gy = q2_const(g, s, ec->Gy, mv);
        const int nn = g.nn;
        const double u = (dmask && dmask[n]) ? 0.0 : gUt[n];
        const double v = (dmask && dmask[n + nn]) ? 0.0 : gUt[n + nn];
        gUs[n] = u + c * gx;
        gUs[n + nn] = v + c * gy;
    }
};

template <class S>
struct AdjFinalQ2 {
    Grid2D g;
    const ElemConstQ2* __restrict__ ec;
    MomCoefQ2T<S> RT;  // transposed rhs operator (rho/dt) M - (1-theta)(rho Ce + mu Le)
    const double *mum, *Uh, *Un, *gUs, *M, *gUns;
    const uint8_t* dmask;
    double *gUn, *gPn;
    double rho, theta, c;  // c = dt / rho
    __device__ void operator()(int n, double (&)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        const int nn = g.nn;
        auto W = [&](int comp, int id) {
            const int i = id + comp * nn;
            const double a = mum[i];
            return (dmask && dmask[i]) ? 0.0 : a;
        };
        double y[2];
        q2_apply<2>(g, s, [&](int a, int e, double (&cf)[9]) { RT.row(a, e, cf); }, W, y);
        double gc[2] = {0.0, 0.0};
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
            const int o = comp * nn;
            conv_grad_q2(g, ec, n, [&](int id) { return W(comp, id); },
                         [&](int id) { return theta * Uh[id + o] + (1.0 - theta) * Un[id + o]; }, 0.5, gc);
        }
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const int i = n + d * nn;
            gUn[i] = y[d] - rho * gc[d] + (gUns ? gUns[i] : 0.0);
        }
        gPn[n] = q2_const<true>(g, s, ec->Gx, [&](int id) { return c * gUs[id] / M[id] - W(0, id); }) +
                 q2_const<true>(g, s, ec->Gy, [&](int id) { return c * gUs[id + nn] / M[id + nn] - W(1, id); });
    }
};

template <class S>
static void ns_vjp_q2_t(Ctx& c, S ce, const double* U_n, double rho, double mu, double dt, const double* gUn_extra,
                        double* gU_n, double* gP_n, int* it) {
    const double theta = 0.5;
    const Grid2D& g = c.g;
    const int nn = g.nn;
    // 6: mu_m = A_m^-T gU*, from the forward x0
    MomCoefQ2T<S> lhsT{g, c.eq2_dev, ce, rho / dt, theta * mu, theta * rho};
    const int maxit = c.maxit_mom > 0 ? c.maxit_mom : 10 * 2 * nn;
    launch_node<2>(c, nn, nullptr,
                   BInitVecQ2<2, MomCoefQ2T<S>>{lhsT, c.a_gus, U_n, c.ud_full, c.dmask, c.a_mx, c.m_r, c.m_rhat,
                                                c.m_p[0], c.m_q[0]},
                   BInitVecFin{c.st, c.tol, maxit});
    launch_node<0>(c, 2 * nn, nullptr, ZeroIfFlag{&c.st->b_zero, c.a_mx}, NoFin{});
    FCFD_CHECK_LAUNCH();
    const double key[4] = {rho, mu, dt, 0.0};
    it[1] = bicgstab_q2_iterate<2, MomCoefQ2T<S>, true>(c, c.g_amom, key, ce.c, lhsT, c.dmask, maxit, c.a_mx,
                                                         c.pred_mom);
    // 7: U_n and P_n cotangents
    MomCoefQ2T<S> RT{g, c.eq2_dev, ce, rho / dt, -(1.0 - theta) * mu, -(1.0 - theta) * rho};
    launch_node<0>(c, nn, nullptr,
                   AdjFinalQ2<S>{g, c.eq2_dev, RT, c.a_mx, c.m_x, U_n, c.a_gus, c.Mlump, gUn_extra, c.dmask, gU_n,
                                 gP_n, rho, theta, dt / rho},
                   NoFin{});
    FCFD_CHECK_LAUNCH();
}

// Requires the forward step's intermediates in the context (U_hat in m_x, Ce(U_n, 1/2) in ce).
// it[0] = pressure adjoint CG iterations, it[1] = momentum adjoint BiCGSTAB iterations.
static void ensure_adj_q2(Ctx& c) {
    if (!c.a_gut) {
        const size_t nn = c.g.nn;
        c.a_sx = dalloc<double>(c, nn);
        c.a_gut = dalloc<double>(c, 2 * nn);
        c.a_guns = dalloc<double>(c, 2 * nn);
        c.a_gus = dalloc<double>(c, 2 * nn);
        c.a_mx = dalloc<double>(c, 2 * nn);
    }
}

void ns_vjp_q2(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* gU,
               const double* gP, const double* gUn_extra, double* gU_n, double* gP_n, int* it) {
    const int nn = c.g.nn;
    ensure_adj_q2(c);
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    const double* gUt = gU;
    if (!gUt) {
        launch_fill(c, c.a_gut, 0.0, 2 * nn);
        gUt = c.a_gut;
    }
    // 3: pressure adjoint rhs
    launch_node<1>(c, nn, nullptr, AdjPrhsQ2{c.g, c.eq2_dev, gUt, c.Mlump, gP, c.dmask, pin, c.c_rhs, dt / rho},
                   MeanFin{&c.st->mean_rhs, nullptr, nn});
    FCFD_CHECK_LAUNCH();
    // 4: CG on the forward operator from the forward x0 (P_n, pins zeroed)
    const int maxit = c.maxit_p > 0 ? c.maxit_p : 1000;
    bool conv;
    it[0] = pcg_q2(c, c.c_rhs, c.set_zeroP ? nullptr : &c.st->mean_rhs, P_n, c.set_zeroP, c.tol, maxit, &conv, true);
    // 5: cotangent of U_hat
    launch_node<0>(c, nn, nullptr, AdjUstarQ2{c.g, c.eq2_dev, gUt, c.c_x, c.dmask, pin, c.st, c.a_gus, rho / dt},
                   NoFin{});
    FCFD_CHECK_LAUNCH();
    // 6, 7
    if (c.ce_format == FCFD_CE_F64)
        ns_vjp_q2_t(c, Ce9F64{(double*)c.ce, c.g.ne}, U_n, rho, mu, dt, gUn_extra, gU_n, gP_n, it);
    else
        ns_vjp_q2_t(c, Ce9F32{(float*)c.ce, c.g.ne}, U_n, rho, mu, dt, gUn_extra, gU_n, gP_n, it);
}

// ------------------------------------------------------------------------------ reverse mode (passive step)
// Plain scalar update (passive_transport.py:39-61): with the forward step's stored element matrices
// A_e, B_e (ade_A, ade_B) and outputs U, phi:
//   1  mu_s = A_s^-T gphi  (BiCGSTAB on the masked transpose, from the forward x0)
//   2  w = mask_phi ? 0 : mu_s;  gphi_n = B_s^T w;  gU~ = gU - 1/2 g_1(w, phi);
//      gU_n,s = -1/2 g_1(w, phi_n)   (A_e, B_e hold +-Ce(U), Ce(U_n) with beta = 1 halved)
//   then the NS-step VJP with gU~ and the extra U_n cotangent gU_n,s.
struct StoredQ2T {
    Grid2D g;
    Ce9F64 m;
    __device__ __forceinline__ void row(int a, int e, double (&cf)[9]) const {
#pragma unroll
        for (int b = 0; b < 9; ++b) cf[b] = m.ld(b * 9 + a, e);
    }
};

// Reverse-mode product of the quad9 SUPG element terms: the quad4 kernel's formulas
// (k_supg_adj_elem, fcfd_scalar.hip) with 9 nodes and 9 Gauss points.  With w the masked scalar
// adjoint, d = phi_n - phi and s = phi_n + phi, per element
//   tau-explicit:  d/dU_n [tau d^T Ce(U_n) w / dt - tau w^T Ct(U_n) s / 2]
//   through tau:   dl/dtau = d^T Ce(U_n) w / dt - w^T Ct s / 2,  tau = S^(-1/2),
//                  S = (|Ct|/|Ce|)^2 + (2/dt)^2 + ((D+1e-6)|Ct|^2 / (u2 |Ce|^2))^2,
//                  |Ce| = |Ce(U)|, |Ct| = |Ct(U_n)|, u2 = mean over the 9 points of |U_n,q|^2
//                  (so du2/dU_n,I = 2/9 sum_q N_I U_n,q).
// The plain convection terms are gathered at the nodes (conv_grad_q2).  Ce(U) and Ct(U_n) are
// formed once per element (81 entries each).  Output gel[((v * 9 + I) * 2 + d) * ne + e], v = 0
// for U_n, 1 for U.
__global__ __launch_bounds__(BLOCK) void k_supg_adj_elem_q2(Grid2D g, const ElemConstQ2* __restrict__ ec,
                                                           const double* __restrict__ Un,
                                                           const double* __restrict__ U,
                                                           const double* __restrict__ mu,
                                                           const uint8_t* __restrict__ mphi,
                                                           const double* __restrict__ phi,
                                                           const double* __restrict__ phin, double D, double dt,
                                                           double* __restrict__ gel) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (e >= g.ne) return;
    const int je = e / g.nex, ie = e - je * g.nex;
    double un[9], vn[9], u[9], v[9], w[9], dl[9], sm[9];
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        const int id = q9_node(g, ie, je, b);
        un[b] = Un[id]; vn[b] = Un[id + g.nn];
        u[b] = U[id]; v[b] = U[id + g.nn];
        const double m = mu[id];
        w[b] = (mphi && mphi[id]) ? 0.0 : m;
        dl[b] = phin[id] - phi[id];
        sm[b] = phin[id] + phi[id];
    }
    double uq[9], vq[9], dq[9], unq[9], vnq[9], dnq[9];
#pragma unroll 1
    for (int q = 0; q < 9; ++q) {
        double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const double Nb = ec->N[q * 9 + b], dx = ec->dNx[q * 9 + b], dy = ec->dNy[q * 9 + b];
            a0 += Nb * u[b]; a1 += Nb * v[b]; a2 += dx * u[b] + dy * v[b];
            b0 += Nb * un[b]; b1 += Nb * vn[b]; b2 += dx * un[b] + dy * vn[b];
        }
        uq[q] = a0; vq[q] = a1; dq[q] = a2;
        unq[q] = b0; vnq[q] = b1; dnq[q] = b2;
    }
    double ce[81], ct[81];
    double nce = 0.0, nct = 0.0, pt = 0.0, u2 = 0.0;
#pragma unroll 1
    for (int a = 0; a < 9; ++a)
#pragma unroll 1
        for (int b = 0; b < 9; ++b) {
            double cu = 0.0, cn = 0.0, tt = 0.0;
            for (int q = 0; q < 9; ++q) {
                const double Na = ec->N[q * 9 + a], Nb = ec->N[q * 9 + b], wt = ec->wq[q];
                const double bx = ec->dNx[q * 9 + b], by = ec->dNy[q * 9 + b];
                cu += wt * Na * (uq[q] * bx + vq[q] * by + dq[q] * Nb);
                cn += wt * Na * (unq[q] * bx + vnq[q] * by + dnq[q] * Nb);
                tt += wt * (unq[q] * ec->dNx[q * 9 + a] + vnq[q] * ec->dNy[q * 9 + a]) * (unq[q] * bx + vnq[q] * by);
            }
            ce[a * 9 + b] = cu;
            ct[a * 9 + b] = tt;
            nce += cu * cu;
            nct += tt * tt;
            pt += dl[a] * cn * w[b] / dt - 0.5 * w[a] * tt * sm[b];
        }
    for (int q = 0; q < 9; ++q) u2 += unq[q] * unq[q] + vnq[q] * vnq[q];
    u2 /= 9.0;
    nce = sqrt(nce);
    nct = sqrt(nct);
    const double Dp = D + 1e-6;
    const double A1 = nct / nce, A3 = Dp * nct * nct / (u2 * nce * nce);
    const double S = A1 * A1 + (2.0 / dt) * (2.0 / dt) + A3 * A3;
    const double tau = 1.0 / sqrt(S);
    const double gS = -0.5 * tau * tau * tau * pt;
    const double k13 = 2.0 * A1 * A1 + 4.0 * A3 * A3;
    const double cG = gS * (-k13 / nce) / nce;  // d l / d Ce(U)_ab = cG Ce_ab
    const double cH = gS * (k13 / nct) / nct;   // d l / d Ct_ab    = cH Ct_ab
    const double gu2 = gS * (-2.0 * A3 * A3 / u2);
    double gn[9][2], gu[9][2];
#pragma unroll
    for (int I = 0; I < 9; ++I) gn[I][0] = gn[I][1] = gu[I][0] = gu[I][1] = 0.0;
#pragma unroll 1
    for (int q = 0; q < 9; ++q) {
        const double* Nq = ec->N + q * 9;
        const double* Xq = ec->dNx + q * 9;
        const double* Yq = ec->dNy + q * 9;
        const double wt = ec->wq[q];
        const double au = unq[q], av = vnq[q];
        double Dq = 0.0, Wq = 0.0, Wx = 0.0, Wy = 0.0, Sx = 0.0, Sy = 0.0;
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            Dq += Nq[c] * dl[c]; Wq += Nq[c] * w[c];
            Wx += Xq[c] * w[c]; Wy += Yq[c] * w[c];
            Sx += Xq[c] * sm[c]; Sy += Yq[c] * sm[c];
        }
        const double aW = au * Wx + av * Wy, aS = au * Sx + av * Sy;
        double gNx = 0.0, gNy = 0.0, gNN = 0.0, hx = 0.0, hy = 0.0;
#pragma unroll 1
        for (int b = 0; b < 9; ++b) {
            double cNa = 0.0, tx = 0.0, ty = 0.0;
#pragma unroll
            for (int a = 0; a < 9; ++a) {
                cNa += ce[a * 9 + b] * Nq[a];
                tx += ct[a * 9 + b] * Xq[a];
                ty += ct[a * 9 + b] * Yq[a];
            }
            gNx += cNa * Xq[b];
            gNy += cNa * Yq[b];
            gNN += cNa * Nq[b];
            const double ab = au * Xq[b] + av * Yq[b];
            hx += tx * ab;
            hy += ty * ab;
        }
#pragma unroll
        for (int I = 0; I < 9; ++I) {
            const double NI = Nq[I];
            gu[I][0] += cG * wt * (NI * gNx + Xq[I] * gNN);
            gu[I][1] += cG * wt * (NI * gNy + Yq[I] * gNN);
            const double cx = (tau / dt) * wt * Dq * (NI * Wx + Xq[I] * Wq);
            const double cy = (tau / dt) * wt * Dq * (NI * Wy + Yq[I] * Wq);
            const double tx = -0.5 * tau * wt * NI * (Wx * aS + Sx * aW);
            const double ty = -0.5 * tau * wt * NI * (Wy * aS + Sy * aW);
            gn[I][0] += cx + tx + 2.0 * cH * wt * NI * hx + (2.0 / 9.0) * gu2 * NI * au;
            gn[I][1] += cy + ty + 2.0 * cH * wt * NI * hy + (2.0 / 9.0) * gu2 * NI * av;
        }
    }
    const size_t ne = g.ne;
#pragma unroll
    for (int I = 0; I < 9; ++I)
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            gel[((size_t)(0 * 9 + I) * 2 + d) * ne + e] = gn[I][d];
            gel[((size_t)(1 * 9 + I) * 2 + d) * ne + e] = gu[I][d];
        }
}

struct AdjScalarQ2 {
    Grid2D g;
    const ElemConstQ2* __restrict__ ec;
    StoredQ2T BT;
    const double *mu, *phi, *phin, *gU;
    const uint8_t* mphi;
    double *gphin, *gUt, *gUns;
    const double* gel;  // SUPG element cotangents (k_supg_adj_elem_q2 layout) or null
    __device__ void operator()(int n, double (&)[1]) const {
        IncQ2 s;
        node_inc_q2(g, n, s);
        auto W = [&](int id) {
            const double a = mu[id];
            return (mphi && mphi[id]) ? 0.0 : a;
        };
        double y[1];
        q2_apply<1>(g, s, [&](int a, int e, double (&cf)[9]) { BT.row(a, e, cf); }, [&](int, int id) { return W(id); },
                    y);
        gphin[n] = y[0];
        double g1[2] = {0.0, 0.0}, g1n[2] = {0.0, 0.0};
        conv_grad_q2(g, ec, n, W, [&](int id) { return phi[id]; }, 1.0, g1);
        conv_grad_q2(g, ec, n, W, [&](int id) { return phin[id]; }, 1.0, g1n);
        double en[2] = {0.0, 0.0}, eu[2] = {0.0, 0.0};
        if (gel) {
            const int j = n / g.nnx, i = n - j * g.nnx;
            const size_t ne = g.ne;
#pragma unroll 1
            for (int k = 0; k < 4; ++k) {
                int ex, ey, I;
                bool has;
                q2_cand(g, i, j, k, ex, ey, I, has);
                if (!has) continue;
                const size_t e = (size_t)ey * g.nex + ex;
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    en[d] += gel[((size_t)(0 * 9 + I) * 2 + d) * ne + e];
                    eu[d] += gel[((size_t)(1 * 9 + I) * 2 + d) * ne + e];
                }
            }
        }
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const int i = n + d * g.nn;
            gUt[i] = (gU ? gU[i] : 0.0) - 0.5 * g1[d] + eu[d];
            gUns[i] = -0.5 * g1n[d] + en[d];
        }
    }
};

// it[0] = scalar adjoint BiCGSTAB, it[1] = pressure adjoint CG, it[2] = momentum adjoint BiCGSTAB
void passive_vjp_q2(Ctx& c, const double* U_n, const double* P_n, const double* phi_n, const double* U,
                    const double* phi, double dt, double rho, double mu, double D, int supg, const double* gU,
                    const double* gP, const double* gphi, double* gU_n, double* gP_n, double* gphi_n, int* it) {
    ensure_adj_q2(c);
    const int nn = c.g.nn;
    const double* gUt = gU;
    const double* gUns = nullptr;
    it[0] = 0;
    if (gphi) {
        StoredQ2T AT{c.g, Ce9F64{c.ade_A, c.g.ne}};
        const int maxit = 10 * nn;
        launch_node<2>(c, nn, nullptr,
                       BInitVecQ2<1, StoredQ2T>{AT, gphi, phi_n, c.phi_ud, c.phi_mask, c.a_sx, c.m_r, c.m_rhat,
                                                c.m_p[0], c.m_q[0]},
                       BInitVecFin{c.st, c.tol, maxit});
        launch_node<0>(c, nn, nullptr, ZeroIfFlag{&c.st->b_zero, c.a_sx}, NoFin{});
        FCFD_CHECK_LAUNCH();
        const double key[4] = {0.0, 0.0, 0.0, 0.0};
        it[0] = bicgstab_q2_iterate<1, StoredQ2T, true>(c, c.g_aphi, key, c.ade_A, AT, c.phi_mask, maxit, c.a_sx,
                                                         c.pred_phi);
        if (supg) {
            if (!c.a_gel) c.a_gel = dalloc<double>(c, 36 * (size_t)c.g.ne);
            k_supg_adj_elem_q2<<<blocks_for(c.g.ne), BLOCK, 0, c.stream>>>(c.g, c.eq2_dev, U_n, U, c.a_sx,
                                                                          c.phi_mask, phi, phi_n, D, dt, c.a_gel);
            FCFD_CHECK_LAUNCH();
        }
        launch_node<0>(c, nn, nullptr,
                       AdjScalarQ2{c.g, c.eq2_dev, StoredQ2T{c.g, Ce9F64{c.ade_B, c.g.ne}}, c.a_sx, phi, phi_n, gU,
                                   c.phi_mask, gphi_n, c.a_gut, c.a_guns, supg ? c.a_gel : nullptr},
                       NoFin{});
        FCFD_CHECK_LAUNCH();
        gUt = c.a_gut;
        gUns = c.a_guns;
    } else {
        launch_fill(c, gphi_n, 0.0, nn);
    }
    int itn[2];
    ns_vjp_q2(c, U_n, P_n, dt, rho, mu, gUt, gP, gUns, gU_n, gP_n, itn);
    it[1] = itn[0];
    it[2] = itn[1];
}

// ------------------------------------------------------------------------------ the step
void four_step_q2(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* U_d,
                  const double* F_p, const double* F, double* U, double* P, int* it) {
    scatter_ud(c, U_d);
    it[0] = solve_momentum_q2(c, U_n, P_n, rho, mu, dt, F, c.m_x);
    launch_node<0>(c, c.g.nn, nullptr,
                   ProjQ2Body{c.g, c.eq2_dev, c.m_x, P_n, c.Mlump, nullptr, nullptr, c.ustar, dt / rho}, NoFin{});
    FCFD_CHECK_LAUNCH();
    it[1] = solve_pressure_q2(c, c.ustar, P_n, F_p, rho, dt, P);
    launch_node<0>(c, c.g.nn, nullptr,
                   ProjQ2Body{c.g, c.eq2_dev, c.ustar, P, c.Mlump, c.ud_full, c.dmask, U, -dt / rho}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

}  // namespace fcfd
