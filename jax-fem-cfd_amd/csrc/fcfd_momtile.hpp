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
template <class S>
struct MomRows {
    double v[4][4];
    __device__ __forceinline__ void load(const S& ce, const Grid2D& g, int i, int j) {
        const Has4 hs = has_at(g, i, j);
        const int e[4] = {hs.h[0] ? (j - 1) * g.nex + (i - 1) : 0, hs.h[1] ? (j - 1) * g.nex + i : 0,
                          hs.h[2] ? j * g.nex + (i - 1) : 0, hs.h[3] ? j * g.nex + i : 0};
#pragma unroll
        for (int k = 0; k < 4; ++k) ce.ld_row(kRow(k), e[k], v[k]);
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
        for (int b = 0; b < 4; ++b) cf[b] = op.cM * op.ec->Me[a * 4 + b] + op.cL * op.ec->Le[a * 4 + b] + op.cC * rw.v[k][b];
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
template <class S, int FC, int FR>
__global__ __launch_bounds__(256) void k_mom_k1t(MomCoef<S> op, const uint8_t* __restrict__ mask,
                                                 const double* __restrict__ r, const double* __restrict__ pold,
                                                 const double* __restrict__ qold, const double* __restrict__ rhat,
                                                 double* __restrict__ pnew, double* __restrict__ qnew,
                                                 const DevState* __restrict__ st, double* __restrict__ partials,
                                                 int pstride, const int* __restrict__ active) {
    static_assert(FC * FR == BLOCK, "one node per thread");
    if (active && *active == 0) return;
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
// s.s, t.s, t.t (BK2S, MT = false).  s is not stored.  TAG only separates the standalone timing.
template <class S, int FC, int FR, int TAG = 0>
__global__ __launch_bounds__(256) void k_mom_k2t(MomCoef<S> op, const uint8_t* __restrict__ mask,
                                                 const double* __restrict__ r, const double* __restrict__ q,
                                                 double* __restrict__ tv, const DevState* __restrict__ st,
                                                 double* __restrict__ partials, int pstride,
                                                 const int* __restrict__ active) {
    static_assert(FC * FR == BLOCK, "one node per thread");
    if (active && *active == 0) return;
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

}  // namespace fcfd
