// Tiled V-cycle kernels of the pinned form on non-periodic quad4 grids (multigrid.py:231-330 with
// the gamma_p rows zeroed, multigrid.py:243-254 and :280-287).
//
// A block owns a rectangle of level nodes and first issues every global load it needs -- the
// windows of r and of the pin flags around its rectangle (and, on the up leg, the window of the
// coarse correction) -- as one batch: the indices are clamped to valid nodes and the values
// outside the grid masked afterwards, so no branch separates the loads and one memory round trip
// serves the whole block.  Everything after that runs out of LDS.
//
// * The pre-smoothed x = 2/3 r / L_diag (0 at the pins) is never stored: the up leg forms it
//   again from r, the value the down leg used bit for bit.
// * L_diag comes from a 3 x 3 table by node class (first / interior / last along each
//   direction): on a uniform grid the node sum of diag(Le) depends only on which elements exist.
// * Every value is formed with the operands and operation order of the node kernels (DownRes,
//   Restrict, Jacobi's pinned path), so the results are bitwise those of the per-level kernels.
#pragma once
#include "fcfd_internal.hpp"
#include "fcfd_ops.hpp"
#include "fcfd_mgnodes.hpp"

namespace fcfd {

// L_diag by node class; v[cj][ci] with c = 0 first, 1 interior, 2 last.  A kernel copies the
// table into LDS (ld_tab_lds) and reads it per node by class index (ld_at): a chain of selects
// between the kernel arguments would become per-lane loads of the argument memory.
template <class T>
struct LdTab {
    T v[3][3];
};
template <class T>
__device__ __forceinline__ void ld_tab_lds(const LdTab<T>& t, T* lds) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) lds[k] = t.v[k / 3][k % 3];
    }
}
template <class T>
__device__ __forceinline__ T ld_at(const T* lds, const Grid2D& g, int i, int j) {
    const int ci = i == 0 ? 0 : (i == g.nnx - 1 ? 2 : 1);
    const int cj = j == 0 ? 0 : (j == g.nny - 1 ? 2 : 1);
    return lds[cj * 3 + ci];
}

// A block's H x W window of a level grid (rows j0 .., columns i0 ..), loaded in registers:
// load() issues every load of the thread (unconditionally, clamped ids), store() hands the
// values to put(a, b, in, v) -- in: the node lies inside the grid.  Separate calls so that the
// loads of several windows are all in flight before the first LDS store.
template <int H, int W, class V>
struct Win {
    static constexpr int N = H * W, NIT = (N + BLOCK - 1) / BLOCK;
    // rows of an LDS array that takes every store of store() (NIT * BLOCK slots, the ones past
    // the window masked): an unconditional store keeps the compiler from sinking the loads of
    // the last, partly filled pass into a branch after the ones it could issue first
    static constexpr int HP = (NIT * BLOCK + W - 1) / W;
    V v[NIT];
    bool in[NIT];
    template <class LD>
    __device__ __forceinline__ void load(const Grid2D& g, int i0, int j0, LD ld) {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int t = threadIdx.x + it * BLOCK;
            const int a = t / W, b = t - a * W;
            const int i = i0 + b, j = j0 + a;
            const bool ok = t < N && i >= 0 && i < g.nnx && j >= 0 && j < g.nny;
            in[it] = ok;
            v[it] = ld(ok ? j * g.nnx + i : 0);
        }
    }
    template <class PUT>
    __device__ __forceinline__ void store(PUT put) const {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int t = threadIdx.x + it * BLOCK;
            const int a = t / W, b = t - a * W;
            put(a, b, in[it], v[it]);
        }
    }
};

template <class TR>
struct RPin {
    TR r;
    uint8_t p;
};

// the pinned pre-smoother from zero: x = 2/3 r / L_diag, 0 at the pins (DownRes, pinned path)
template <class T>
__device__ __forceinline__ T presmooth(T r, T ld, bool pinned) {
    const T x = (T(2.0 / 3.0) * r) / ld;
    return pinned ? T(0) : x;
}

// The incident elements of node (i, j) of a non-periodic grid: lower-left, lower-right,
// upper-left, upper-right (NodeNbr::has, from the node's indices instead of its id)
struct Has4 {
    bool h[4];
};
__device__ __forceinline__ Has4 has_at(const Grid2D& g, int i, int j) {
    const bool l = i >= 1, r = i < g.nex, d = j >= 1, u = j < g.ney;
    return Has4{{l && d, r && d, l && u, r && u}};
}
// ebe_const with the incident-element flags given
template <class T>
__device__ __forceinline__ T ebe_has(const Has4& hs, const double* __restrict__ A, const T (&X)[3][3]) {
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = kRow(k);
        T t = T(0);
#pragma unroll
        for (int b = 0; b < 4; ++b) t += T(A[a * 4 + b]) * X[kCy(k, b)][kCx(k, b)];
        acc += hs.h[k] ? t : T(0);
    }
    return acc;
}

// the same with A^T (the divergence uses G_i^T; ebe_const_T's order)
template <class T>
__device__ __forceinline__ T ebe_has_T(const Has4& hs, const double* __restrict__ A, const T (&X)[3][3]) {
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = kRow(k);
        T t = T(0);
#pragma unroll
        for (int b = 0; b < 4; ++b) t += T(A[b * 4 + a]) * X[kCy(k, b)][kCx(k, b)];
        acc += hs.h[k] ? t : T(0);
    }
    return acc;
}

// f = r - L x at node (i, j) of level grid g (pinned rows 0) from the x stencil X
template <class T>
__device__ __forceinline__ T pinned_residual(const Grid2D& g, int i, int j, const double* Le, T r, bool pinned,
                                             const T (&X)[3][3]) {
    const T rh = pinned ? T(0) : r;
    const T lx = pinned ? T(0) : ebe_has(has_at(g, i, j), Le, X);
    return rh - lx;
}

// R = I^T (unscaled) of a fine residual window at coarse node (cc, rr) of the window
template <class T, class FV>
__device__ __forceinline__ T restrict_at(FV val) {
    const T E = ((val(1, 0) + val(-1, 0)) + val(0, 1)) + val(0, -1);
    const T D = ((val(1, 1) + val(1, -1)) + val(-1, 1)) + val(-1, -1);
    return (val(0, 0) + T(0.5) * E) + T(0.25) * D;
}

// interp_at (bilinear prolongation, multigrid.py:90-140) at fine node (i, j) of a non-periodic
// grid, coarse values read through yv(J, I) (absolute coarse indices; the two columns / rows
// coincide for even fine indices, where 0.5 a + 0.5 a = a exactly).  The same operands and order
// as the per-level kernels' interp_at and interp3x3, so the same bits.
template <class T, class YV>
__device__ __forceinline__ T interp_pt(int i, int j, YV yv) {
    const int i0 = i >> 1, i1 = (i + 1) >> 1, j0 = j >> 1, j1 = (j + 1) >> 1;
    const T a = yv(j0, i0), b = yv(j0, i1), c = yv(j1, i0), d = yv(j1, i1);
    const T r0 = T(0.5) * a + T(0.5) * b;
    const T r1 = T(0.5) * c + T(0.5) * d;
    return T(0.5) * r0 + T(0.5) * r1;
}

// The damped Jacobi of the pinned up leg at node (i, j) (Jacobi's pinned path) from the stencil
// values vv(p, q) = x + I y at stencil point (p, q) = (dy + 1, dx + 1), 0 where the neighbour is
// missing or pinned; r, L_diag and the pin flag of the node itself.
template <class T, class VV>
__device__ __forceinline__ T pinned_jacobi_at(const Grid2D& g, int i, int j, const double* Leoff, T r, T ld,
                                              bool pinned, VV vv) {
    T X[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) X[a][b] = vv(a, b);
    const T off = ebe_has(has_at(g, i, j), Leoff, X);
    T v = (T(OMEGA) * (r - off)) / ld + T(1.0 - OMEGA) * X[1][1];
    if (pinned) v = T(0);
    return v;
}

// ------------------------------------------------------------------------------ down leg, one level
// A block owns CX x CY coarse nodes (2 CX x 2 CY fine ones).  It loads r and the pins of the
// fine window rows 2 J0 - 2 .. 2 J0 + 2 CY, forms x there, f = r - L x on the window one in, and
// restricts f into its coarse nodes.
template <class T, class TR, int CX, int CY, bool EARLY>
__global__ __launch_bounds__(256) void k_mgt_down(Grid2D gf, Grid2D gc, const double* __restrict__ Le, LdTab<T> ldt_,
                                                  const TR* __restrict__ r, const uint8_t* __restrict__ pin,
                                                  T* __restrict__ rc, const int* __restrict__ active) {
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    constexpr int XW = 2 * CX + 3, XR = 2 * CY + 3;
    constexpr int FW = 2 * CX + 1, FR = 2 * CY + 1;
    __shared__ T ldl[9];
    constexpr int XP = Win<XR, XW, RPin<TR>>::HP;
    __shared__ T xs[XP][XW];
    __shared__ T rs[XP][XW];
    __shared__ uint8_t ps[XP][XW];
    __shared__ T fs[FR][FW];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int J0 = (lb / nbx) * CY, I0 = (lb - (lb / nbx) * nbx) * CX;
    const int fi0 = 2 * I0 - 2, fj0 = 2 * J0 - 2;
    Win<XR, XW, RPin<TR>> w;
    w.load(gf, fi0, fj0, [&](int k) { return RPin<TR>{r[k], pin[k]}; });
    const int go = gate_read(active);
    ld_tab_lds(ldt_, ldl);
    __syncthreads();
    if (!gate_open(go)) return;
    w.store([&](int a, int b, bool in, const RPin<TR>& v) {
        const T rv = (T)v.r;
        const bool pn = v.p != 0;
        xs[a][b] = in ? presmooth(rv, ld_at(ldl, gf, fi0 + b, fj0 + a), pn) : T(0);
        rs[a][b] = rv;
        ps[a][b] = pn;
    });
    __syncthreads();
    for (int t = threadIdx.x; t < FR * FW; t += 256) {
        const int a = t / FW, b = t - a * FW;
        const int fi = fi0 + 1 + b, fj = fj0 + 1 + a;
        T fv = T(0);
        if (fi >= 0 && fi < gf.nnx && fj >= 0 && fj < gf.nny) {
            T X[3][3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) X[dy][dx] = xs[a + dy][b + dx];
            fv = pinned_residual(gf, fi, fj, Le, rs[a + 1][b + 1], ps[a + 1][b + 1] != 0, X);
        }
        fs[a][b] = fv;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < CY * CX; t += 256) {
        const int jr = t / CX, ci = t - jr * CX;
        const int I = I0 + ci, J = J0 + jr;
        if (I < gc.nnx && J < gc.nny) {
            const int cc = 2 * ci + 1, rr = 2 * jr + 1;
            rc[J * gc.nnx + I] = restrict_at<T>([&](int dy, int dx) -> T { return fs[rr + dy][cc + dx]; });
        }
    }
}

// The same first down leg with the PCG's x / r update fused in (pinned form, level 0, inside an
// iteration: CgK2's pass, jax_iterative_solvers.py:108-109): the window holds r_new = r - alpha Ap
// (CgK2's operation, alpha from the device state), the block stores r_new of its own fine nodes
// into r_out (r is double-buffered: neighbouring blocks still read r at their halo) and
// x += alpha p there, and writes its partial of r_new . r_new into partials[slot * pstride + lb].
template <class T, int CX, int CY, bool EARLY>
__global__ __launch_bounds__(256) void k_mgt_down_cg(Grid2D gf, Grid2D gc, const double* __restrict__ Le,
                                                     LdTab<T> ldt_, const double* __restrict__ r,
                                                     const double* __restrict__ Ap, const double* __restrict__ p,
                                                     double* __restrict__ xcg, double* __restrict__ r_out,
                                                     const uint8_t* __restrict__ pin, T* __restrict__ rc,
                                                     const DevState* __restrict__ st, double* __restrict__ partials,
                                                     int pstride, int slot, const int* __restrict__ active) {
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    constexpr int XW = 2 * CX + 3, XR = 2 * CY + 3;
    constexpr int FW = 2 * CX + 1, FR = 2 * CY + 1;
    constexpr int OW = 2 * CX, OR = 2 * CY, NO = (OW * OR + BLOCK - 1) / BLOCK;  // owned fine nodes
    __shared__ T ldl[9];
    struct RAP {
        double r, ap;
        uint8_t p;
    };
    constexpr int XP = Win<XR, XW, RAP>::HP;
    __shared__ T xs[XP][XW];
    __shared__ T rs[XP][XW];
    __shared__ double rd[OR][OW];
    __shared__ uint8_t ps[XP][XW];
    __shared__ T fs[FR][FW];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int J0 = (lb / nbx) * CY, I0 = (lb - (lb / nbx) * nbx) * CX;
    const int fi0 = 2 * I0 - 2, fj0 = 2 * J0 - 2;
    const double alpha = st->c_alpha;
    Win<XR, XW, RAP> w;
    w.load(gf, fi0, fj0, [&](int k) { return RAP{r[k], Ap[k], pin[k]}; });
    double po[NO], xo[NO];
#pragma unroll
    for (int m = 0; m < NO; ++m) {
        const int t = threadIdx.x + m * BLOCK;
        const int i = 2 * I0 + t % OW, j = 2 * J0 + t / OW;
        const int n = (t < OW * OR && i < gf.nnx && j < gf.nny) ? j * gf.nnx + i : 0;
        po[m] = p[n];
        xo[m] = xcg[n];
    }
    const int go = gate_read(active);
    ld_tab_lds(ldt_, ldl);
    __syncthreads();
    if (!gate_open(go)) return;
    w.store([&](int a, int b, bool in, const RAP& v) {
        const double rk = v.r - alpha * v.ap;
        const T rv = (T)rk;
        const bool pn = v.p != 0;
        xs[a][b] = in ? presmooth(rv, ld_at(ldl, gf, fi0 + b, fj0 + a), pn) : T(0);
        rs[a][b] = rv;
        ps[a][b] = pn;
        if (a >= 2 && a < 2 + OR && b >= 2 && b < 2 + OW) rd[a - 2][b - 2] = rk;
    });
    __syncthreads();
    double acc[1] = {0.0};
#pragma unroll
    for (int m = 0; m < NO; ++m) {
        const int t = threadIdx.x + m * BLOCK;
        const int oi = t % OW, oj = t / OW;
        const int i = 2 * I0 + oi, j = 2 * J0 + oj;
        if (t < OW * OR && i < gf.nnx && j < gf.nny) {
            const int n = j * gf.nnx + i;
            const double rk = rd[oj][oi];
            r_out[n] = rk;
            xcg[n] = xo[m] + alpha * po[m];
            acc[0] += rk * rk;
        }
    }
    for (int t = threadIdx.x; t < FR * FW; t += 256) {
        const int a = t / FW, b = t - a * FW;
        const int fi = fi0 + 1 + b, fj = fj0 + 1 + a;
        T fv = T(0);
        if (fi >= 0 && fi < gf.nnx && fj >= 0 && fj < gf.nny) {
            T X[3][3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) X[dy][dx] = xs[a + dy][b + dx];
            fv = pinned_residual(gf, fi, fj, Le, rs[a + 1][b + 1], ps[a + 1][b + 1] != 0, X);
        }
        fs[a][b] = fv;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < CY * CX; t += 256) {
        const int jr = t / CX, ci = t - jr * CX;
        const int I = I0 + ci, J = J0 + jr;
        if (I < gc.nnx && J < gc.nny) {
            const int cc = 2 * ci + 1, rr = 2 * jr + 1;
            rc[J * gc.nnx + I] = restrict_at<T>([&](int dy, int dx) -> T { return fs[rr + dy][cc + dx]; });
        }
    }
    reduce_epilogue<1>(acc, partials + (size_t)slot * pstride, pstride, lb);
}

// ------------------------------------------------------------------------------ down leg, two levels
// Small levels l -> m = l + 1 -> c = l + 2: a block owns M2_R x M2_C nodes of level c and forms,
// in LDS, what they depend on: f_m on the (2 M2_R + 1) x (2 M2_C + 1) m-nodes around them, hence
// r_m and x_m on a one-wider m-window, r_m being the restriction of f_l on the (2 (2 M2_R + 3) + 1)-
// row l-window, hence x_l on a one-wider l-window.  It stores r_m and r_c of the nodes it owns.
constexpr int M2_R = 2, M2_C = 16;
template <class T, class TR, bool EARLY>
__global__ __launch_bounds__(256) void k_mgt_down2(Grid2D gf, Grid2D gm, Grid2D gc, const double* __restrict__ Lef,
                                                   const double* __restrict__ Lem, LdTab<T> ldf_, LdTab<T> ldm_,
                                                   const TR* __restrict__ rf, const uint8_t* __restrict__ pinf,
                                                   const uint8_t* __restrict__ pinm, T* __restrict__ rm,
                                                   T* __restrict__ rc, const int* __restrict__ active) {
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    constexpr int XRm = 2 * M2_R + 3, XWm = 2 * M2_C + 3;  // m-window of x_m, r_m
    constexpr int FRm = 2 * M2_R + 1, FWm = 2 * M2_C + 1;  // m-window of f_m
    constexpr int FRl = 2 * XRm + 1, FWl = 2 * XWm + 1;    // l-window of f_l
    constexpr int XRl = FRl + 2, XWl = FWl + 2;            // l-window of x_l
    __shared__ T ldfl[9], ldml[9];
    constexpr int XPl = Win<XRl, XWl, RPin<TR>>::HP, XPm = Win<XRm, XWm, uint8_t>::HP;
    __shared__ T xl[XPl][XWl];
    __shared__ T rl[XPl][XWl];
    __shared__ uint8_t pl[XPl][XWl];
    __shared__ T fl[FRl][FWl];
    __shared__ T xms[XRm][XWm];
    __shared__ T rms[XRm][XWm];
    __shared__ uint8_t pms[XPm][XWm];
    __shared__ T fm[FRm][FWm];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int J0 = (lb / nbx) * M2_R, I0 = (lb - (lb / nbx) * nbx) * M2_C;
    const int mj0 = 2 * J0 - 2, mi0 = 2 * I0 - 2;    // origin of the x_m window
    const int lj0 = 2 * mj0 - 2, li0 = 2 * mi0 - 2;  // origin of the x_l window (f_l: one in)
    Win<XRl, XWl, RPin<TR>> w;
    w.load(gf, li0, lj0, [&](int k) { return RPin<TR>{rf[k], pinf[k]}; });
    Win<XRm, XWm, uint8_t> wp;
    wp.load(gm, mi0, mj0, [&](int k) { return pinm[k]; });
    const int go = gate_read(active);
    ld_tab_lds(ldf_, ldfl);
    ld_tab_lds(ldm_, ldml);
    __syncthreads();
    if (!gate_open(go)) return;
    w.store([&](int a, int b, bool in, const RPin<TR>& v) {
        const T rv = (T)v.r;
        const bool pn = v.p != 0;
        xl[a][b] = in ? presmooth(rv, ld_at(ldfl, gf, li0 + b, lj0 + a), pn) : T(0);
        rl[a][b] = rv;
        pl[a][b] = pn;
    });
    wp.store([&](int a, int b, bool, uint8_t p) { pms[a][b] = p; });
    __syncthreads();
    // f_l = r_l - L x_l (pinned rows 0), 0 outside the grid
    for (int t = threadIdx.x; t < FRl * FWl; t += 256) {
        const int a = t / FWl, b = t - a * FWl;
        const int fi = li0 + 1 + b, fj = lj0 + 1 + a;
        T fv = T(0);
        if (fi >= 0 && fi < gf.nnx && fj >= 0 && fj < gf.nny) {
            T X[3][3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) X[dy][dx] = xl[a + dy][b + dx];
            fv = pinned_residual(gf, fi, fj, Lef, rl[a + 1][b + 1], pl[a + 1][b + 1] != 0, X);
        }
        fl[a][b] = fv;
    }
    __syncthreads();
    // r_m = R f_l on the x_m window, x_m from it; the owned r_m stored
    for (int t = threadIdx.x; t < XRm * XWm; t += 256) {
        const int a = t / XWm, b = t - a * XWm;
        const int mi = mi0 + b, mj = mj0 + a;
        const bool in = mi >= 0 && mi < gm.nnx && mj >= 0 && mj < gm.nny;
        T rv = T(0), xv = T(0);
        if (in) {
            const int rr = 2 * a + 1, cc = 2 * b + 1;
            rv = restrict_at<T>([&](int dy, int dx) -> T { return fl[rr + dy][cc + dx]; });
            xv = presmooth(rv, ld_at(ldml, gm, mi, mj), pms[a][b] != 0);
            if (a >= 2 && a < 2 + 2 * M2_R && b >= 2 && b < 2 + 2 * M2_C) rm[mj * gm.nnx + mi] = rv;
        }
        rms[a][b] = rv;
        xms[a][b] = xv;
    }
    __syncthreads();
    // f_m = r_m - L x_m (pinned rows 0), 0 outside the grid
    for (int t = threadIdx.x; t < FRm * FWm; t += 256) {
        const int a = t / FWm, b = t - a * FWm;
        const int mi = mi0 + 1 + b, mj = mj0 + 1 + a;
        T fv = T(0);
        if (mi >= 0 && mi < gm.nnx && mj >= 0 && mj < gm.nny) {
            T X[3][3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) X[dy][dx] = xms[a + dy][b + dx];
            fv = pinned_residual(gm, mi, mj, Lem, rms[a + 1][b + 1], pms[a + 1][b + 1] != 0, X);
        }
        fm[a][b] = fv;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < M2_R * M2_C; t += 256) {
        const int jr = t / M2_C, ci = t - jr * M2_C;
        const int I = I0 + ci, J = J0 + jr;
        if (I < gc.nnx && J < gc.nny) {
            const int cc = 2 * ci + 1, rr = 2 * jr + 1;
            rc[J * gc.nnx + I] = restrict_at<T>([&](int dy, int dx) -> T { return fm[rr + dy][cc + dx]; });
        }
    }
}

// ------------------------------------------------------------------------------ up leg, one level
// A block owns FC x FR nodes of level l: it loads r and the pins on the window one wider and
// the coarse correction y_{l+1} on the window its prolongation reads, forms v = x + I y at every
// window node once (0 where missing or pinned), then one damped Jacobi per owned node from v.
// ZM = 0: y_l stored (T).  ZM = 1 (level 0, before the first PCG iteration: CgGamma0's pass):
// z = y stored into znew, p_old cleared, block partial of r.z.  ZM = 2 (level 0 inside an
// iteration: CgK3's pass): z stored, partials of r.z and (flexible form) r.z_old.  ZT: the type
// z is stored in -- FP64, or the FP32 V-cycle's own type where the CG's reader widens it on load
// (the FP64 value an FP64 store of the widened y would hold, at half the bytes).
template <class T, class TR, int FC, int FR, int ZM, class ZT, bool EARLY>
__global__ __launch_bounds__(256) void k_mgt_up(Grid2D gf, Grid2D gc, const double* __restrict__ Leoff, LdTab<T> ldt_,
                                                const TR* __restrict__ r, const uint8_t* __restrict__ pin,
                                                const T* __restrict__ yc, T* __restrict__ y, ZT* __restrict__ znew,
                                                const ZT* __restrict__ zold, double* __restrict__ pzero, int flex,
                                                double* __restrict__ partials, int pstride,
                                                const int* __restrict__ active) {
    static_assert(FC % 2 == 0 && FR % 2 == 0 && (FC * FR) % BLOCK == 0,
                  "even tile sides (the coarse window starts at tile origin / 2 - 1), whole nodes per thread");
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    constexpr int WW = FC + 2, WR = FR + 2;          // fine window (stencil of the owned nodes)
    constexpr int CW = FC / 2 + 3, CR = FR / 2 + 3;  // coarse window (prolongation sources)
    constexpr int NPT = FC * FR / BLOCK;
    __shared__ T ldl[9];
    constexpr int WP = Win<WR, WW, RPin<TR>>::HP, CP = Win<CR, CW, T>::HP;
    __shared__ T rs[WP][WW];
    __shared__ uint8_t ps[WP][WW];  // 1 pinned, 2 outside the grid
    __shared__ T vs[WR][WW];
    __shared__ T ys[CP][CW];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * FR, i0 = (lb - (lb / nbx) * nbx) * FC;
    const int wi0 = i0 - 1, wj0 = j0 - 1;
    const int ci0 = (i0 >> 1) - 1, cj0 = (j0 >> 1) - 1;
    Win<WR, WW, RPin<TR>> w;
    w.load(gf, wi0, wj0, [&](int k) { return RPin<TR>{r[k], pin[k]}; });
    Win<CR, CW, T> wc;
    wc.load(gc, ci0, cj0, [&](int k) { return yc[k]; });
    // the owned nodes' FP64 r and z_old for the CG's dot products (level 0)
    double rown[NPT], zown[NPT];
    if constexpr (ZM != 0) {
#pragma unroll
        for (int m = 0; m < NPT; ++m) {
            const int t = threadIdx.x + m * BLOCK;
            const int i = i0 + t % FC, j = j0 + t / FC;
            const int n = (i < gf.nnx && j < gf.nny) ? j * gf.nnx + i : 0;
            rown[m] = (double)r[n];
            zown[m] = (ZM == 2 && flex) ? (double)zold[n] : 0.0;
        }
    }
    const int go = gate_read(active);
    ld_tab_lds(ldt_, ldl);
    w.store([&](int a, int b, bool in, const RPin<TR>& v) {
        rs[a][b] = (T)v.r;
        ps[a][b] = in ? (v.p != 0 ? 1 : 0) : 2;
    });
    wc.store([&](int a, int b, bool, T v) { ys[a][b] = v; });
    __syncthreads();
    if (!gate_open(go)) return;
    for (int t = threadIdx.x; t < WR * WW; t += BLOCK) {
        const int a = t / WW, b = t - a * WW;
        const int fi = wi0 + b, fj = wj0 + a;
        const T x = (T(2.0 / 3.0) * rs[a][b]) / ld_at(ldl, gf, fi, fj);
        const T iy = interp_pt<T>(fi, fj, [&](int J, int I) { return ys[J - cj0][I - ci0]; });
        vs[a][b] = ps[a][b] == 0 ? x + iy : T(0);
    }
    __syncthreads();
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < NPT; ++m) {
        const int t = threadIdx.x + m * BLOCK;
        const int bi = t % FC, bj = t / FC;
        const int i = i0 + bi, j = j0 + bj;
        if (i < gf.nnx && j < gf.nny) {
            const int n = j * gf.nnx + i;
            const T v = pinned_jacobi_at(gf, i, j, Leoff, rs[bj + 1][bi + 1], ld_at(ldl, gf, i, j),
                                         ps[bj + 1][bi + 1] != 0, [&](int p, int q) { return vs[bj + p][bi + q]; });
            if constexpr (ZM == 0) {
                y[n] = v;
            } else {
                const double zk = (double)v;
                znew[n] = (ZT)v;
                if constexpr (ZM == 1) pzero[n] = 0.0;
                acc[0] += rown[m] * zk;
                if constexpr (ZM == 2) {
                    if (flex) acc[1] += rown[m] * zown[m];
                }
            }
        }
    }
    if constexpr (ZM != 0) {
        constexpr int NR = ZM == 1 ? 1 : 2;
        double a2[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) a2[q] = acc[q];
        reduce_epilogue<NR>(a2, partials, pstride, lb);
    }
}

// ------------------------------------------------------------------------------ up leg, two levels
// Small levels c = l + 2 -> m = l + 1 -> l: a block owns a U2_R x U2_C tile of level-l nodes.  It
// loads r and the pins of level l on the tile's stencil window, r and the pins of level m around
// the m-nodes the tile's prolongation reads, and y_c on the window their own prolongation reads;
// then, in LDS: v_m = x_m + I y_c on the m-window, the level-m Jacobi y_m (not stored),
// v_l = x_l + I y_m on the l-window and the level-l Jacobi of its nodes.
constexpr int U2_R = 8, U2_C = 32;
template <class T, class TR, bool EARLY>
__global__ __launch_bounds__(256) void k_mgt_up2(Grid2D gf, Grid2D gm, Grid2D gc, const double* __restrict__ Leoff_f,
                                                 const double* __restrict__ Leoff_m, LdTab<T> ldf_, LdTab<T> ldm_,
                                                 const TR* __restrict__ rf, const uint8_t* __restrict__ pinf,
                                                 const T* __restrict__ rm, const uint8_t* __restrict__ pinm,
                                                 const T* __restrict__ yc, T* __restrict__ yf,
                                                 const int* __restrict__ active) {
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    constexpr int LR = U2_R + 2, LC = U2_C + 2;            // level-l window (tile stencil)
    constexpr int MR = U2_R / 2 + 3, MC = U2_C / 2 + 3;    // m-nodes the l-window's prolongation reads
    constexpr int MXR = MR + 2, MXC = MC + 2;              // their stencil window
    constexpr int CRW = MXR / 2 + 3, CCW = MXC / 2 + 3;    // coarse window of the m-window's prolongation
    __shared__ T ldfl[9], ldml[9];
    constexpr int LP = Win<LR, LC, RPin<TR>>::HP, MP = Win<MXR, MXC, RPin<T>>::HP, CP = Win<CRW, CCW, T>::HP;
    __shared__ T rl[LP][LC];
    __shared__ uint8_t pl[LP][LC];  // 1 pinned, 2 outside the grid
    __shared__ T vl[LR][LC];
    __shared__ T rms[MP][MXC];
    __shared__ uint8_t pms[MP][MXC];
    __shared__ T vm[MXR][MXC];
    __shared__ T ycs[CP][CCW];
    __shared__ T ym[MR][MC];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * U2_R, i0 = (lb - (lb / nbx) * nbx) * U2_C;
    const int li0 = i0 - 1, lj0 = j0 - 1;
    const int mj0 = (j0 - 1) >> 1, mi0 = (i0 - 1) >> 1;  // first m-node of the l-window's prolongation
    const int xmj0 = mj0 - 1, xmi0 = mi0 - 1;              // its stencil window
    const int cj0 = (xmj0 >> 1) - 1, ci0 = (xmi0 >> 1) - 1;
    Win<LR, LC, RPin<TR>> wl;
    wl.load(gf, li0, lj0, [&](int k) { return RPin<TR>{rf[k], pinf[k]}; });
    Win<MXR, MXC, RPin<T>> wm;
    wm.load(gm, xmi0, xmj0, [&](int k) { return RPin<T>{rm[k], pinm[k]}; });
    Win<CRW, CCW, T> wc;
    wc.load(gc, ci0, cj0, [&](int k) { return yc[k]; });
    const int go = gate_read(active);
    ld_tab_lds(ldf_, ldfl);
    ld_tab_lds(ldm_, ldml);
    wl.store([&](int a, int b, bool in, const RPin<TR>& v) {
        rl[a][b] = (T)v.r;
        pl[a][b] = in ? (v.p != 0 ? 1 : 0) : 2;
    });
    wm.store([&](int a, int b, bool in, const RPin<T>& v) {
        rms[a][b] = v.r;
        pms[a][b] = in ? (v.p != 0 ? 1 : 0) : 2;
    });
    wc.store([&](int a, int b, bool, T v) { ycs[a][b] = v; });
    __syncthreads();
    if (!gate_open(go)) return;
    for (int t = threadIdx.x; t < MXR * MXC; t += BLOCK) {
        const int a = t / MXC, b = t - a * MXC;
        const int mi = xmi0 + b, mj = xmj0 + a;
        const T x = (T(2.0 / 3.0) * rms[a][b]) / ld_at(ldml, gm, mi, mj);
        const T iy = interp_pt<T>(mi, mj, [&](int J, int I) { return ycs[J - cj0][I - ci0]; });
        vm[a][b] = pms[a][b] == 0 ? x + iy : T(0);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < MR * MC; t += BLOCK) {
        const int a = t / MC, b = t - a * MC;
        const int mi = mi0 + b, mj = mj0 + a;
        T v = T(0);
        if (pms[a + 1][b + 1] != 2)
            v = pinned_jacobi_at(gm, mi, mj, Leoff_m, rms[a + 1][b + 1], ld_at(ldml, gm, mi, mj),
                                 pms[a + 1][b + 1] != 0, [&](int p, int q) { return vm[a + p][b + q]; });
        ym[a][b] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < LR * LC; t += BLOCK) {
        const int a = t / LC, b = t - a * LC;
        const int fi = li0 + b, fj = lj0 + a;
        const T x = (T(2.0 / 3.0) * rl[a][b]) / ld_at(ldfl, gf, fi, fj);
        const T iy = interp_pt<T>(fi, fj, [&](int J, int I) { return ym[J - mj0][I - mi0]; });
        vl[a][b] = pl[a][b] == 0 ? x + iy : T(0);
    }
    __syncthreads();
    const int bj = threadIdx.x / U2_C, bi = threadIdx.x % U2_C;
    const int j = j0 + bj, i = i0 + bi;
    if (i < gf.nnx && j < gf.nny)
        yf[j * gf.nnx + i] = pinned_jacobi_at(gf, i, j, Leoff_f, rl[bj + 1][bi + 1], ld_at(ldfl, gf, i, j),
                                              pl[bj + 1][bi + 1] != 0,
                                              [&](int p, int q) { return vl[bj + p][bi + q]; });
}

// ------------------------------------------------------------------------------ up leg, three levels
// Levels c = l + 3 -> n = l + 2 -> m = l + 1 -> l in one kernel, for an l too large for the
// two-level kernel's next pairing (the level below the two-level pairs): a block owns a U3_R x U3_C
// tile of level-l nodes and recomputes, in LDS, everything its Jacobi depends on -- v_n = x_n +
// I y_c on the n-window and the level-n Jacobi y_n on the n-nodes the m-window's prolongation
// reads, then v_m and y_m the same way, then v_l and the level-l Jacobi of its own nodes.  Every
// value is formed by the operands and order of the one-level kernel, so y_l is bitwise the chain
// of three one-level up legs.
constexpr int U3_R = 8, U3_C = 32;
template <class T, bool EARLY>
__global__ __launch_bounds__(256) void k_mgt_up3(Grid2D gf, Grid2D gm, Grid2D gn, Grid2D gc,
                                                 const double* __restrict__ Leoff_f, const double* __restrict__ Leoff_m,
                                                 const double* __restrict__ Leoff_n, LdTab<T> ldf_, LdTab<T> ldm_,
                                                 LdTab<T> ldn_, const T* __restrict__ rf, const uint8_t* __restrict__ pinf,
                                                 const T* __restrict__ rm, const uint8_t* __restrict__ pinm,
                                                 const T* __restrict__ rn, const uint8_t* __restrict__ pinn,
                                                 const T* __restrict__ yc, T* __restrict__ yf,
                                                 const int* __restrict__ active) {
    static_assert(U3_R * U3_C == BLOCK && U3_R % 2 == 0 && U3_C % 2 == 0,
                  "one node per thread; even tile sides (the windows start at tile origin / 2 - 1)");
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    constexpr int LR = U3_R + 2, LC = U3_C + 2;          // level-l window (tile stencil)
    constexpr int MR = U3_R / 2 + 3, MC = U3_C / 2 + 3;  // m-nodes the l-window's prolongation reads
    constexpr int MXR = MR + 2, MXC = MC + 2;            // their stencil window
    constexpr int NR = MXR / 2 + 3, NC = MXC / 2 + 3;    // n-nodes the m-window's prolongation reads
    constexpr int NXR = NR + 2, NXC = NC + 2;            // their stencil window
    constexpr int CRW = NXR / 2 + 3, CCW = NXC / 2 + 3;  // coarse window of the n-window's prolongation
    __shared__ T ldfl[9], ldml[9], ldnl[9];
    constexpr int LP = Win<LR, LC, RPin<T>>::HP, MP = Win<MXR, MXC, RPin<T>>::HP;
    constexpr int NP = Win<NXR, NXC, RPin<T>>::HP, CP = Win<CRW, CCW, T>::HP;
    __shared__ T rl[LP][LC];
    __shared__ uint8_t pl[LP][LC];  // 1 pinned, 2 outside the grid
    __shared__ T vl[LR][LC];
    __shared__ T rms[MP][MXC];
    __shared__ uint8_t pms[MP][MXC];
    __shared__ T vm[MXR][MXC];
    __shared__ T ym[MR][MC];
    __shared__ T rns[NP][NXC];
    __shared__ uint8_t pns[NP][NXC];
    __shared__ T vn[NXR][NXC];
    __shared__ T yn[NR][NC];
    __shared__ T ycs[CP][CCW];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * U3_R, i0 = (lb - (lb / nbx) * nbx) * U3_C;
    const int li0 = i0 - 1, lj0 = j0 - 1;
    const int mj0 = (j0 - 1) >> 1, mi0 = (i0 - 1) >> 1;  // first m-node of the l-window's prolongation
    const int xmj0 = mj0 - 1, xmi0 = mi0 - 1;              // its stencil window
    const int nj0 = (xmj0 >> 1) - 1, ni0 = (xmi0 >> 1) - 1;  // first n-node of the m-window's prolongation
    const int xnj0 = nj0 - 1, xni0 = ni0 - 1;
    const int cj0 = (xnj0 >> 1) - 1, ci0 = (xni0 >> 1) - 1;
    Win<LR, LC, RPin<T>> wl;
    wl.load(gf, li0, lj0, [&](int k) { return RPin<T>{rf[k], pinf[k]}; });
    Win<MXR, MXC, RPin<T>> wm;
    wm.load(gm, xmi0, xmj0, [&](int k) { return RPin<T>{rm[k], pinm[k]}; });
    Win<NXR, NXC, RPin<T>> wn;
    wn.load(gn, xni0, xnj0, [&](int k) { return RPin<T>{rn[k], pinn[k]}; });
    Win<CRW, CCW, T> wc;
    wc.load(gc, ci0, cj0, [&](int k) { return yc[k]; });
    const int go = gate_read(active);
    ld_tab_lds(ldf_, ldfl);
    ld_tab_lds(ldm_, ldml);
    ld_tab_lds(ldn_, ldnl);
    wl.store([&](int a, int b, bool in, const RPin<T>& v) {
        rl[a][b] = v.r;
        pl[a][b] = in ? (v.p != 0 ? 1 : 0) : 2;
    });
    wm.store([&](int a, int b, bool in, const RPin<T>& v) {
        rms[a][b] = v.r;
        pms[a][b] = in ? (v.p != 0 ? 1 : 0) : 2;
    });
    wn.store([&](int a, int b, bool in, const RPin<T>& v) {
        rns[a][b] = v.r;
        pns[a][b] = in ? (v.p != 0 ? 1 : 0) : 2;
    });
    wc.store([&](int a, int b, bool, T v) { ycs[a][b] = v; });
    __syncthreads();
    if (!gate_open(go)) return;
    for (int t = threadIdx.x; t < NXR * NXC; t += BLOCK) {
        const int a = t / NXC, b = t - a * NXC;
        const int ni = xni0 + b, nj = xnj0 + a;
        const T x = (T(2.0 / 3.0) * rns[a][b]) / ld_at(ldnl, gn, ni, nj);
        const T iy = interp_pt<T>(ni, nj, [&](int J, int I) { return ycs[J - cj0][I - ci0]; });
        vn[a][b] = pns[a][b] == 0 ? x + iy : T(0);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NR * NC; t += BLOCK) {
        const int a = t / NC, b = t - a * NC;
        const int ni = ni0 + b, nj = nj0 + a;
        T v = T(0);
        if (pns[a + 1][b + 1] != 2)
            v = pinned_jacobi_at(gn, ni, nj, Leoff_n, rns[a + 1][b + 1], ld_at(ldnl, gn, ni, nj),
                                 pns[a + 1][b + 1] != 0, [&](int p, int q) { return vn[a + p][b + q]; });
        yn[a][b] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < MXR * MXC; t += BLOCK) {
        const int a = t / MXC, b = t - a * MXC;
        const int mi = xmi0 + b, mj = xmj0 + a;
        const T x = (T(2.0 / 3.0) * rms[a][b]) / ld_at(ldml, gm, mi, mj);
        const T iy = interp_pt<T>(mi, mj, [&](int J, int I) { return yn[J - nj0][I - ni0]; });
        vm[a][b] = pms[a][b] == 0 ? x + iy : T(0);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < MR * MC; t += BLOCK) {
        const int a = t / MC, b = t - a * MC;
        const int mi = mi0 + b, mj = mj0 + a;
        T v = T(0);
        if (pms[a + 1][b + 1] != 2)
            v = pinned_jacobi_at(gm, mi, mj, Leoff_m, rms[a + 1][b + 1], ld_at(ldml, gm, mi, mj),
                                 pms[a + 1][b + 1] != 0, [&](int p, int q) { return vm[a + p][b + q]; });
        ym[a][b] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < LR * LC; t += BLOCK) {
        const int a = t / LC, b = t - a * LC;
        const int fi = li0 + b, fj = lj0 + a;
        const T x = (T(2.0 / 3.0) * rl[a][b]) / ld_at(ldfl, gf, fi, fj);
        const T iy = interp_pt<T>(fi, fj, [&](int J, int I) { return ym[J - mj0][I - mi0]; });
        vl[a][b] = pl[a][b] == 0 ? x + iy : T(0);
    }
    __syncthreads();
    const int bj = threadIdx.x / U3_C, bi = threadIdx.x % U3_C;
    const int j = j0 + bj, i = i0 + bi;
    if (i < gf.nnx && j < gf.nny)
        yf[j * gf.nnx + i] = pinned_jacobi_at(gf, i, j, Leoff_f, rl[bj + 1][bi + 1], ld_at(ldfl, gf, i, j),
                                              pl[bj + 1][bi + 1] != 0,
                                              [&](int p, int q) { return vl[bj + p][bi + q]; });
}

// ------------------------------------------------------------------------------ PCG K1, tiled
// p = z + beta p_old (CgK1S's staged value), Ap = L p (pinned rows 0), block partial of p . Ap
// (jax_iterative_solvers.py:105-107), for a pinned non-periodic grid: a block owns FC x FR
// nodes, loads z, p_old and the pins on the window one wider in one batch.  ZT: the type z is
// stored in (the V-cycle's), widened to FP64 on load.
template <int FC, int FR, class ZT, bool EARLY>
__global__ __launch_bounds__(256) void k_cg_k1t(Grid2D g, const double* __restrict__ Le, const uint8_t* __restrict__ pin,
                                                const ZT* __restrict__ z, const double* __restrict__ pold,
                                                double* __restrict__ pnew, double* __restrict__ Ap,
                                                const DevState* __restrict__ st, double* __restrict__ partials,
                                                int pstride, const int* __restrict__ active) {
    static_assert((FC * FR) % BLOCK == 0, "whole nodes per thread");
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    constexpr int WW = FC + 2, WR = FR + 2, NPT = FC * FR / BLOCK;
    struct ZP {
        double z, p;
        uint8_t pin;
    };  // z widened on load
    constexpr int WP = Win<WR, WW, ZP>::HP;
    __shared__ double ps[WP][WW];
    __shared__ uint8_t pns[WP][WW];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * FR, i0 = (lb - (lb / nbx) * nbx) * FC;
    const double beta = st->c_beta;
    Win<WR, WW, ZP> w;
    w.load(g, i0 - 1, j0 - 1, [&](int k) { return ZP{(double)z[k], pold[k], pin[k]}; });
    const int go = gate_read(active);
    w.store([&](int a, int b, bool in, const ZP& v) {
        ps[a][b] = in ? v.z + beta * v.p : 0.0;
        pns[a][b] = v.pin;
    });
    __syncthreads();
    if (!gate_open(go)) return;
    double acc[1] = {0.0};
#pragma unroll
    for (int m = 0; m < NPT; ++m) {
        const int t = threadIdx.x + m * BLOCK;
        const int bi = t % FC, bj = t / FC;
        const int i = i0 + bi, j = j0 + bj;
        if (i < g.nnx && j < g.nny) {
            const int n = j * g.nnx + i;
            double X[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) X[a][b] = ps[bj + a][bi + b];
            const double av = pns[bj + 1][bi + 1] ? 0.0 : ebe_has(has_at(g, i, j), Le, X);
            pnew[n] = X[1][1];
            Ap[n] = av;
            acc[0] += X[1][1] * av;
        }
    }
    reduce_epilogue<1>(acc, partials, pstride, lb);
}

// ------------------------------------------------------------------------------ U* + PCG init, tiled
// The projection U* = U_hat + (dt/rho) G P_n / M (UstarBody, standard.py:155) and the pinned PCG
// init with the pressure rhs formed in it (CgInitD: b = (rho/dt)(D U* - F_p), 0 at the pins;
// x0 = P_n with the pins zeroed; r = b - L x0; partials b.b, r.r, sum r) in one pass, for a
// non-periodic grid.  A block owns FC x FR nodes: it loads P_n on the window two wider (G P_n at
// every node around its own) and U_hat, the lumped mass (one value per node: both components
// carry the same) and the pins on the window one wider, forms U* there in LDS (stored for its own
// nodes: the final projection reads it), then the divergence and the Laplacian of its nodes.
// The operands and order are those of UstarBody and CgInitD.
template <int FC, int FR>
__global__ __launch_bounds__(256) void k_cginit_us(Grid2D g, const ElemConst* __restrict__ ec,
                                                   const double* __restrict__ Le, const double* __restrict__ Uh,
                                                   const double* __restrict__ Pn, const double* __restrict__ M,
                                                   const double* __restrict__ Fp, const uint8_t* __restrict__ pin,
                                                   double cus, double cb, double* __restrict__ Us,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ partials, int pstride,
                                                   const int* __restrict__ gate) {
    static_assert(FC * FR == BLOCK, "one node per thread");
    constexpr int PW = FC + 4, PR = FR + 4;  // P_n window: two wider
    constexpr int UW = FC + 2, UR = FR + 2;  // U_hat / M / pins / U* window: one wider
    struct UMP {
        double u0, u1, m;
        uint8_t p;
    };
    constexpr int PP = Win<PR, PW, double>::HP, UP = Win<UR, UW, UMP>::HP;
    __shared__ double ps[PP][PW];
    __shared__ double uhs[2][UP][UW];
    __shared__ double mss[UP][UW];
    __shared__ uint8_t pns[UP][UW];
    __shared__ double uss[2][UR][UW];
    const int nn = g.nn;
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * FR, i0 = (lb - (lb / nbx) * nbx) * FC;
    const int bi = threadIdx.x % FC, bj = threadIdx.x / FC;
    const int i = i0 + bi, j = j0 + bj;
    const bool own = i < g.nnx && j < g.nny;
    const int n = own ? j * g.nnx + i : 0;
    Win<PR, PW, double> wp;
    wp.load(g, i0 - 2, j0 - 2, [&](int k) { return Pn[k]; });
    Win<UR, UW, UMP> wu;
    wu.load(g, i0 - 1, j0 - 1, [&](int k) { return UMP{Uh[k], Uh[k + nn], M[k], pin[k]}; });
    const double fpo = Fp ? Fp[n] : 0.0;
    const int go = gate_read(gate);
    wp.store([&](int a, int b, bool in, double v) { ps[a][b] = in ? v : 0.0; });
    wu.store([&](int a, int b, bool in, const UMP& v) {
        uhs[0][a][b] = v.u0;
        uhs[1][a][b] = v.u1;
        mss[a][b] = v.m;
        pns[a][b] = in && v.p != 0;
    });
    __syncthreads();
    if (!gate_open(go)) return;
    for (int t = threadIdx.x; t < UR * UW; t += BLOCK) {
        const int a = t / UW, b = t - a * UW;
        const int ui = i0 - 1 + b, uj = j0 - 1 + a;
        double v0 = 0.0, v1 = 0.0;
        if (ui >= 0 && ui < g.nnx && uj >= 0 && uj < g.nny) {
            double X[3][3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) X[dy][dx] = ps[a + dy][b + dx];
            const Has4 hs = has_at(g, ui, uj);
            const double gx = ebe_has(hs, ec->Gx, X), gy = ebe_has(hs, ec->Gy, X);
            v0 = uhs[0][a][b] + (cus * gx) / mss[a][b];
            v1 = uhs[1][a][b] + (cus * gy) / mss[a][b];
        }
        uss[0][a][b] = v0;
        uss[1][a][b] = v1;
    }
    __syncthreads();
    double acc[3] = {0.0, 0.0, 0.0};
    if (own) {
        double Xu[3][3], Xv[3][3], X0[3][3];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                Xu[dy][dx] = uss[0][bj + dy][bi + dx];
                Xv[dy][dx] = uss[1][bj + dy][bi + dx];
                X0[dy][dx] = pns[bj + dy][bi + dx] ? 0.0 : ps[bj + 1 + dy][bi + 1 + dx];
            }
        Us[n] = Xu[1][1];
        Us[n + nn] = Xv[1][1];
        const Has4 hs = has_at(g, i, j);
        const bool pn = pns[bj + 1][bi + 1] != 0;
        const double div = ebe_has_T(hs, ec->Gx, Xu) + ebe_has_T(hs, ec->Gy, Xv);
        double bn = cb * (div - fpo);
        if (pn) bn = 0.0;
        const double ax = pn ? 0.0 : ebe_has(hs, Le, X0);
        const double rn = bn - ax;
        x[n] = X0[1][1];
        r[n] = rn;
        acc[0] += bn * bn;
        acc[1] += rn * rn;
        acc[2] += rn;
    }
    reduce_epilogue<3>(acc, partials, pstride, lb);
}

// ------------------------------------------------------------------------------ output + projection, tiled
// The pinned form's pressure output (PoutBody: P = the solution with the pins zeroed) and the
// final projection U = U* - (dt/rho) G P / M with the Dirichlet values restored (FinalBody,
// standard.py:159-163) in one pass, for a non-periodic grid: a block owns FC x FR nodes, stages
// the solution x with the pins zeroed on the window one wider -- that is P -- stores P for its
// own nodes and forms G P from the window.  The operands and order are those of PoutBody and
// FinalBody, so P and U are bitwise theirs; P is written only here, so it may alias P_n.
template <int FC, int FR>
__global__ __launch_bounds__(256) void k_final_t(Grid2D g, const ElemConst* __restrict__ ec,
                                                 const double* __restrict__ x, const uint8_t* __restrict__ pin,
                                                 const double* __restrict__ Us, const double* __restrict__ M,
                                                 const double* __restrict__ ud, const uint8_t* __restrict__ mask,
                                                 double c, double* __restrict__ P, double* __restrict__ U,
                                                 const int* __restrict__ gate) {
    static_assert(FC * FR == BLOCK, "one node per thread");
    if (gate && *gate == 0) return;
    constexpr int WW = FC + 2, WR = FR + 2;
    struct XP {
        double x;
        uint8_t p;
    };
    constexpr int WP = Win<WR, WW, XP>::HP;
    __shared__ double ps[WP][WW];
    const int nn = g.nn;
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * FR, i0 = (lb - (lb / nbx) * nbx) * FC;
    const int bi = threadIdx.x % FC, bj = threadIdx.x / FC;
    const int i = i0 + bi, j = j0 + bj;
    const bool own = i < g.nnx && j < g.nny;
    const int n = own ? j * g.nnx + i : 0;
    Win<WR, WW, XP> w;
    w.load(g, i0 - 1, j0 - 1, [&](int k) { return XP{x[k], pin[k]}; });
    const double us0 = Us[n], us1 = Us[n + nn], m0 = M[n], m1 = M[n + nn];
    const double ud0 = ud[n], ud1 = ud[n + nn];
    const uint8_t k0 = mask[n], k1 = mask[n + nn];
    w.store([&](int a, int b, bool in, const XP& v) { ps[a][b] = in ? (v.p ? 0.0 : v.x) : 0.0; });
    __syncthreads();
    if (own) {
        double X[3][3];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) X[dy][dx] = ps[bj + dy][bi + dx];
        P[n] = X[1][1];
        const Has4 hs = has_at(g, i, j);
        const double gx = ebe_has(hs, ec->Gx, X), gy = ebe_has(hs, ec->Gy, X);
        double u = us0 - (c * gx) / m0;
        double v = us1 - (c * gy) / m1;
        if (k0) u = ud0;
        if (k1) v = ud1;
        U[n] = u;
        U[n + nn] = v;
    }
}

}  // namespace fcfd
