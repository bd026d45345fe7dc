// Tiled BiCGSTAB operator kernels of the 2D momentum solve on non-periodic quad4 grids
// (jax_fem_cfd/utils/jax_iterative_solvers.py:123-167 on the operator of standard.py:14-43).
//
// A block owns FC x FR nodes (one per thread).  Before its first barrier every thread issues all
// of its global loads in one batch: the window of the staged vector (both components, one node
// wider than the tile), its node's four element-matrix rows and its Dirichlet flags.  The values
// are those of the LDS-staged node kernels BK1S / BK2S (op_apply's operands and order, the same
// masked rows), so the iterates are bitwise the same; only the block partials of the dots are
// grouped by tile.
// Included by fcfd_bicgstab.hpp after MomCoef and bicg_s.
#pragma once
#include "fcfd_mgtile.hpp"

namespace fcfd {

template <class LOp>
struct IsMomCoef : std::false_type {};
template <class S>
struct IsMomCoef<MomCoef<S>> : std::true_type {};


// The four rows a node owns of its incident elements (op.row of MomCoef: cM Me + cL Le + cC Ce_e),
// loaded before any of them is used; missing elements read element 0 and are masked by has.
__device__ __forceinline__ void inc_elems(const Grid2D& g, int i, int j, int (&e)[4]) {
    const Has4 hs = has_at(g, i, j);
    e[0] = hs.h[0] ? (j - 1) * g.nex + (i - 1) : 0;
    e[1] = hs.h[1] ? (j - 1) * g.nex + i : 0;
    e[2] = hs.h[2] ? j * g.nex + (i - 1) : 0;
    e[3] = hs.h[3] ? j * g.nex + i : 0;
}
template <class S>
struct MomRows {
    double v[4][4];
    __device__ __forceinline__ void load(const S& ce, const Grid2D& g, int i, int j) {
        int e[4];
        inc_elems(g, i, j, e);
#pragma unroll
        for (int k = 0; k < 4; ++k) ce.ld_row(kRow(k), e[k], v[k]);
    }
    __device__ __forceinline__ double at(int k, int b) const { return v[k][b]; }
};
// FP32 storage: the rows stay FP32 in registers until they are used (half the registers across
// the barrier); the value used is the same (double)(float) conversion CeF32::ld_row makes
template <>
struct MomRows<CeF32> {
    float4 v[4];
    __device__ __forceinline__ void load(const CeF32& ce, const Grid2D& g, int i, int j) {
        int e[4];
        inc_elems(g, i, j, e);
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = *reinterpret_cast<const float4*>(ce.c + ce_idx(kRow(k) * 4, e[k], ce.ne));
    }
    __device__ __forceinline__ double at(int k, int b) const {
        return (double)(b == 0 ? v[k].x : (b == 1 ? v[k].y : (b == 2 ? v[k].z : v[k].w)));
    }
};

// y = op_apply<2>(op, s, X, y) with the rows given
template <class S>
__device__ __forceinline__ void mom_apply(const MomCoef<S>& op, const Has4& hs, const MomRows<S>& rw,
                                          const double (&X)[2][3][3], double (&y)[2]) {
    y[0] = 0.0;
    y[1] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = kRow(k);
        double cf[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) cf[b] = op.cM * op.ec->Me[a * 4 + b] + op.cL * op.ec->Le[a * 4 + b] + op.cC * rw.at(k, b);
        double t[2] = {0.0, 0.0};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
#pragma unroll
            for (int c = 0; c < 2; ++c) t[c] += cf[b] * X[c][kCy(k, b)][kCx(k, b)];
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) y[c] += hs.h[k] ? t[c] : 0.0;
    }
}

struct Val2M {
    double v0, v1;
};

// ------------------------------------------------------------------------------ K1
// p = r + beta (p_old - omega q_old) (p = r at the first iteration), q = A p with the Dirichlet
// rows replaced by p, block partial of rhat . q (BK1S, MT = false).
template <class S, int FC, int FR, bool EARLY>
__global__ __launch_bounds__(256) void k_mom_k1t(MomCoef<S> op, const uint8_t* __restrict__ mask,
                                                 const double* __restrict__ r, const double* __restrict__ pold,
                                                 const double* __restrict__ qold, const double* __restrict__ rhat,
                                                 double* __restrict__ pnew, double* __restrict__ qnew,
                                                 const DevState* __restrict__ st, double* __restrict__ partials,
                                                 int pstride, const int* __restrict__ active) {
    static_assert(FC * FR == BLOCK, "one node per thread");
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    const int go = gate_read(active);
    constexpr int WW = FC + 2, WR = FR + 2;
    constexpr int WP = Win<WR, WW, Val2M>::HP;
    __shared__ double xs[2][WP][WW];
    const Grid2D& g = op.g;
    const int nn = g.nn;
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * FR, i0 = (lb - (lb / nbx) * nbx) * FC;
    const int bi = threadIdx.x % FC, bj = threadIdx.x / FC;
    const int i = i0 + bi, j = j0 + bj;
    const bool own = i < g.nnx && j < g.nny;
    const int n = own ? j * g.nnx + i : 0;
    MomRows<S> rw;
    rw.load(op.ce, g, own ? i : 0, own ? j : 0);
    const uint8_t m0 = mask[n], m1 = mask[n + nn];
    const double rh0 = rhat[n], rh1 = rhat[n + nn];
    if (st->b_k == 0) {
        Win<WR, WW, Val2M> w;
        w.load(g, i0 - 1, j0 - 1, [&](int k) { return Val2M{r[k], r[k + nn]}; });
        w.store([&](int a, int b, bool in, const Val2M& v) {
            xs[0][a][b] = in ? v.v0 : 0.0;
            xs[1][a][b] = in ? v.v1 : 0.0;
        });
    } else {
        const double beta = st->b_beta, omega = st->b_omega;
        struct V6 {
            double r0, p0, q0, r1, p1, q1;
        };
        Win<WR, WW, V6> w;
        w.load(g, i0 - 1, j0 - 1,
               [&](int k) { return V6{r[k], pold[k], qold[k], r[k + nn], pold[k + nn], qold[k + nn]}; });
        w.store([&](int a, int b, bool in, const V6& v) {
            xs[0][a][b] = in ? v.r0 + beta * (v.p0 - omega * v.q0) : 0.0;
            xs[1][a][b] = in ? v.r1 + beta * (v.p1 - omega * v.q1) : 0.0;
        });
    }
    __syncthreads();
    if (!gate_open(go)) return;
    double acc[1] = {0.0};
    if (own) {
        double X[2][3][3];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) X[c][a][b] = xs[c][bj + a][bi + b];
        double y[2];
        mom_apply(op, has_at(g, i, j), rw, X, y);
        const bool mk[2] = {m0 != 0, m1 != 0};
        const double rh[2] = {rh0, rh1};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double pc = X[c][1][1];
            const double yc = mk[c] ? pc : y[c];
            pnew[n + c * nn] = pc;
            qnew[n + c * nn] = yc;
            acc[0] += rh[c] * yc;
        }
    }
    reduce_epilogue<1>(acc, partials, pstride, lb);
}

// ------------------------------------------------------------------------------ K2
// s = r - alpha q (bicg_s), t = A s with the Dirichlet rows replaced by s, block partials of
// s.s, t.s, t.t (BK2S, MT = false).  s is not stored.  TAG = 1: the standalone
// timing copy; TAG = 2: an in-solver launch that also tests its gate at entry (EARLY).
template <class S, int FC, int FR, int TAG = 0>
__global__ __launch_bounds__(256) void k_mom_k2t(MomCoef<S> op, const uint8_t* __restrict__ mask,
                                                 const double* __restrict__ r, const double* __restrict__ q,
                                                 double* __restrict__ tv, const DevState* __restrict__ st,
                                                 double* __restrict__ partials, int pstride,
                                                 const int* __restrict__ active) {
    static_assert(FC * FR == BLOCK, "one node per thread");
    if constexpr (TAG == 2) {
        if (*active == 0) return;
    }
    const int go = gate_read(active);
    constexpr int WW = FC + 2, WR = FR + 2;
    struct V4 {
        double r0, q0, r1, q1;
    };
    constexpr int WP = Win<WR, WW, V4>::HP;
    __shared__ double xs[2][WP][WW];
    const Grid2D& g = op.g;
    const int nn = g.nn;
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * FR, i0 = (lb - (lb / nbx) * nbx) * FC;
    const int bi = threadIdx.x % FC, bj = threadIdx.x / FC;
    const int i = i0 + bi, j = j0 + bj;
    const bool own = i < g.nnx && j < g.nny;
    const int n = own ? j * g.nnx + i : 0;
    const double alpha = st->b_alpha;
    Win<WR, WW, V4> w;
    w.load(g, i0 - 1, j0 - 1, [&](int k) { return V4{r[k], q[k], r[k + nn], q[k + nn]}; });
    MomRows<S> rw;
    rw.load(op.ce, g, own ? i : 0, own ? j : 0);
    const uint8_t m0 = mask[n], m1 = mask[n + nn];
    w.store([&](int a, int b, bool in, const V4& v) {
        xs[0][a][b] = in ? bicg_s(v.r0, alpha, v.q0) : 0.0;
        xs[1][a][b] = in ? bicg_s(v.r1, alpha, v.q1) : 0.0;
    });
    __syncthreads();
    if (!gate_open(go)) return;
    double acc[3] = {0.0, 0.0, 0.0};
    if (own) {
        double X[2][3][3];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) X[c][a][b] = xs[c][bj + a][bi + b];
        double y[2];
        mom_apply(op, has_at(g, i, j), rw, X, y);
        const bool mk[2] = {m0 != 0, m1 != 0};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double sc = X[c][1][1];
            const double yc = mk[c] ? sc : y[c];
            tv[n + c * nn] = yc;
            acc[0] += sc * sc;
            acc[1] += yc * sc;
            acc[2] += yc * yc;
        }
    }
    reduce_epilogue<3>(acc, partials, pstride, lb);
}

// ------------------------------------------------------------------------------ init
// The momentum solve's init (BInitMomS, standard.py:24-39): rhs b = (rho/dt) M_c U_n
// - (1 - theta)(rho C + mu L) U_n - G P_n + F with the Dirichlet rows set to U_d, x0 = U_n with
// the Dirichlet values, r0 = rhat = b - A x0; partials b.b, r0.r0.  The window holds U_n, x0
// (both components) and P_n; each element row feeds both the lhs and the rhs coefficients.
template <class S, int FC, int FR>
__global__ __launch_bounds__(256) void k_mom_initt(MomCoef<S> lhs, MomCoef<S> rhs, const ElemConst* __restrict__ ec,
                                                   const double* __restrict__ Pn, const double* __restrict__ F,
                                                   const double* __restrict__ xn, const double* __restrict__ ud,
                                                   const uint8_t* __restrict__ mask, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ rhat,
                                                   double* __restrict__ partials, int pstride,
                                                   const int* __restrict__ gate) {
    static_assert(FC * FR == BLOCK, "one node per thread");
    if (gate && *gate == 0) return;
    constexpr int WW = FC + 2, WR = FR + 2;
    struct V7 {
        double x0, x1, u0, u1, p;
        uint8_t m0, m1;
    };
    constexpr int WP = Win<WR, WW, V7>::HP;
    __shared__ double xs[2][WP][WW];   // U_n
    __shared__ double x0s[2][WP][WW];  // x0
    __shared__ double pps[WP][WW];     // P_n
    __shared__ double uds[2][WP][WW];  // U_d
    __shared__ uint8_t ms[2][WP][WW];  // Dirichlet flags
    const Grid2D& g = lhs.g;
    const int nn = g.nn;
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * FR, i0 = (lb - (lb / nbx) * nbx) * FC;
    const int bi = threadIdx.x % FC, bj = threadIdx.x / FC;
    const int i = i0 + bi, j = j0 + bj;
    const bool own = i < g.nnx && j < g.nny;
    const int n = own ? j * g.nnx + i : 0;
    Win<WR, WW, V7> w;
    w.load(g, i0 - 1, j0 - 1, [&](int k) {
        return V7{xn[k], xn[k + nn], ud[k], ud[k + nn], Pn[k], mask[k], mask[k + nn]};
    });
    MomRows<S> rw;
    rw.load(lhs.ce, g, own ? i : 0, own ? j : 0);
    double Fo[2] = {0.0, 0.0};
    if (F) {
        Fo[0] = F[n];
        Fo[1] = F[n + nn];
    }
    w.store([&](int a, int b, bool in, const V7& v) {
        xs[0][a][b] = in ? v.x0 : 0.0;
        xs[1][a][b] = in ? v.x1 : 0.0;
        x0s[0][a][b] = in ? (v.m0 ? v.u0 : v.x0) : 0.0;
        x0s[1][a][b] = in ? (v.m1 ? v.u1 : v.x1) : 0.0;
        pps[a][b] = in ? v.p : 0.0;
        uds[0][a][b] = v.u0;
        uds[1][a][b] = v.u1;
        ms[0][a][b] = v.m0;
        ms[1][a][b] = v.m1;
    });
    __syncthreads();
    double acc[2] = {0.0, 0.0};
    if (own) {
        const Has4 hs = has_at(g, i, j);
        double b[2] = {0.0, 0.0}, ax[2] = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int a = kRow(k);
            double cl[4], cr[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                cl[q] = lhs.cM * lhs.ec->Me[a * 4 + q] + lhs.cL * lhs.ec->Le[a * 4 + q] + lhs.cC * rw.at(k, q);
                cr[q] = rhs.cM * rhs.ec->Me[a * 4 + q] + rhs.cL * rhs.ec->Le[a * 4 + q] + rhs.cC * rw.at(k, q);
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                double tl = 0.0, tr = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int dy = kCy(k, q), dx = kCx(k, q);
                    tl += cl[q] * x0s[c][bj + dy][bi + dx];
                    tr += cr[q] * xs[c][bj + dy][bi + dx];
                }
                ax[c] += hs.h[k] ? tl : 0.0;
                b[c] += hs.h[k] ? tr : 0.0;
            }
        }
        double gx = 0.0, gy = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int a = kRow(k);
            double tx = 0.0, ty = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double p = pps[bj + kCy(k, q)][bi + kCx(k, q)];
                tx += ec->Gx[a * 4 + q] * p;
                ty += ec->Gy[a * 4 + q] * p;
            }
            gx += hs.h[k] ? tx : 0.0;
            gy += hs.h[k] ? ty : 0.0;
        }
        b[0] -= gx;
        b[1] -= gy;
        if (F) {
            b[0] += Fo[0];
            b[1] += Fo[1];
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int ii = n + c * nn;
            const double x0 = x0s[c][bj + 1][bi + 1];
            const bool m = ms[c][bj + 1][bi + 1] != 0;
            const double bc = m ? uds[c][bj + 1][bi + 1] : b[c];
            const double av = m ? x0 : ax[c];
            const double rc = bc - av;
            x[ii] = x0;
            r[ii] = rc;
            rhat[ii] = rc;
            acc[0] += bc * bc;
            acc[1] += rc * rc;
        }
    }
    reduce_epilogue<2>(acc, partials, pstride, lb);
}

}  // namespace fcfd
