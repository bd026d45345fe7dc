// Multigrid pieces that act on node grids only (no element operator): the Jacobi weight,
// the Neumann-form pre-smoother from zero, restriction R = I^T, bilinear prolongation and the
// Neumann coarse Jacobi.  Shared by the quad4 and quad9 V-cycles (multigrid.py:90-330): a level's
// Grid2D is its node grid, whatever the element.
#pragma once
#include "fcfd_internal.hpp"
#include "fcfd_ops.hpp"

namespace fcfd {

constexpr double OMEGA = 2.0 / 3.0;

// xh = 2/3 (r - mean r) / L_diag   (multigrid.py:257-261), sum -> mean(xh)
template <class T, class TR>
struct DownX {
    const TR* r; const T* Ld; T* x; const DevState* st; int l;
    __device__ void operator()(int n, double (&acc)[1]) const {
        const T mr = (T)st->mg[l][MG_R];
        const T v = (T(2.0 / 3.0) * ((T)r[n] - mr)) / Ld[n];
        x[n] = v;
        acc[0] += (double)v;
    }
};

// coarse[I] = R (f - mean f): R = I^T without the 1/4 factor (multigrid.py:142-212);
// missing neighbours (domain edges) contribute zero.  Periodic grids restrict the padded fine
// grid with the same formulas and drop the wrap row/column, so only negative fine indices can
// be missing there.
template <class T>
struct Restrict {
    Grid2D gf, gc; const T* f; T* rc; const DevState* st; int l; int neumann;
    __device__ void operator()(int nc, double (&acc)[1]) const {
        const int I = nc / gc.nnx, J = nc - I * gc.nnx;
        const int fi = 2 * I, fj = 2 * J;
        const T mf = neumann ? (T)st->mg[l][MG_F] : T(0);
        auto val = [&](int dy, int dx) -> T {
            int a = fi + dy, b = fj + dx;
            bool ok = a >= 0 && b >= 0 && (gf.periodic || (a < gf.nny && b < gf.nnx));
            return ok ? f[a * gf.nnx + b] - mf : T(0);
        };
        const T E = ((val(1, 0) + val(-1, 0)) + val(0, 1)) + val(0, -1);
        const T D = ((val(1, 1) + val(1, -1)) + val(-1, 1)) + val(-1, -1);
        const T v = (val(0, 0) + T(0.5) * E) + T(0.25) * D;
        rc[nc] = v;
        acc[0] += (double)v;
    }
};

// bilinear prolongation of the coarse vector (minus its mean) at fine node n
// (multigrid.py:90-140; periodic coarse grids wrap)
template <class T>
__device__ __forceinline__ T interp_at(const Grid2D& gf, const Grid2D& gc, const T* yc, T my, int n) {
    // Branch-free (lanes of a wave alternate parity): the two coarse columns / rows coincide for
    // even fine indices, where 0.5 a + 0.5 a = a exactly.
    const int j = n / gf.nnx, i = n - j * gf.nnx;
    const int i0 = i >> 1, j0 = j >> 1;
    int i1 = (i + 1) >> 1, j1 = (j + 1) >> 1;
    if (gc.periodic) {
        i1 = i1 >= gc.nnx ? i1 - gc.nnx : i1;
        j1 = j1 >= gc.nny ? j1 - gc.nny : j1;
    }
    const T a = yc[j0 * gc.nnx + i0] - my, b = yc[j0 * gc.nnx + i1] - my;
    const T c = yc[j1 * gc.nnx + i0] - my, d = yc[j1 * gc.nnx + i1] - my;
    const T r0 = T(0.5) * a + T(0.5) * b;
    const T r1 = T(0.5) * c + T(0.5) * d;
    return T(0.5) * r0 + T(0.5) * r1;
}

template <class T>
struct Prolong {
    Grid2D gf, gc; const T* yc; T* p; const DevState* st; int l;
    __device__ void operator()(int n, double (&acc)[1]) const {
        const T v = interp_at(gf, gc, yc, (T)st->mg[l + 1][MG_Y], n);
        p[n] = v;
        acc[0] += (double)v;
    }
};

// Neumann: x = omega (rN - Le_off 0) / L_diag + (1 - omega) 0, then its mean
template <class T, class TR>
__global__ __launch_bounds__(1024) void k_coarse_jacobi(int nn, const T* Ld, const TR* r, T* x, DevState* st, int l,
                                                        const int* active) {
    if (active && *active == 0) return;
    const T mr = (T)st->mg[l][MG_R];
    double acc[1] = {0.0};
    for (int n = threadIdx.x; n < nn; n += blockDim.x) {
        const T v = (T(OMEGA) * (((T)r[n] - mr) - T(0))) / Ld[n] + T(1.0 - OMEGA) * T(0);
        x[n] = v;
        acc[0] += (double)v;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) st->mg[l][MG_Y] = acc[0] / nn;
}

// interp_at at all 9 stencil points of a node from one 3x3 patch of coarse values: fine index f
// interpolates coarse (f >> 1, (f + 1) >> 1), so fine f - 1 .. f + 1 touch coarse
// floor((f - 1) / 2) + {0, 1, 2}.  Which patch entries pair up depends only on the parity of the
// node's (i, j): selected per lane, with the same operands and operation order as interp_at, so
// every value is bitwise the one interp_at returns.  9 coarse loads instead of 36.
// yv(cy, cx): the coarse value at coarse row cy, column cx (global memory, or a block's LDS copy)
template <class T, class YV>
__device__ __forceinline__ void interp3x3_v(const Grid2D& gc, YV yv, T my, const NodeNbr& s, T (&Y)[3][3]) {
    const int bx = (s.i - 1) >> 1, by = (s.j - 1) >> 1;  // arithmetic shift: floor
    int cx[3], cy[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        int x = bx + t, y = by + t;
        if (gc.periodic) {
            x = x < 0 ? x + gc.nnx : (x >= gc.nnx ? x - gc.nnx : x);
            y = y < 0 ? y + gc.nny : (y >= gc.nny ? y - gc.nny : y);
        } else {
            x = x < 0 ? 0 : (x >= gc.nnx ? gc.nnx - 1 : x);
            y = y < 0 ? 0 : (y >= gc.nny ? gc.nny - 1 : y);
        }
        cx[t] = x;
        cy[t] = y;
    }
    T P[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) P[a][b] = yv(cy[a], cx[b]) - my;
    // pair (first, second) of patch columns for fine offsets -1, 0, +1:
    //   even i: (0,1) (1,1) (1,2)      odd i: (0,0) (0,1) (1,1)
    const bool ox = s.i & 1, oy = s.j & 1;
    T H[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T p0 = P[a][0], p1 = P[a][1], p2 = P[a][2];
        H[a][0] = T(0.5) * p0 + T(0.5) * (ox ? p0 : p1);
        H[a][1] = T(0.5) * (ox ? p0 : p1) + T(0.5) * p1;
        H[a][2] = T(0.5) * p1 + T(0.5) * (ox ? p1 : p2);
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const T h0 = H[0][b], h1 = H[1][b], h2 = H[2][b];
        Y[0][b] = T(0.5) * h0 + T(0.5) * (oy ? h0 : h1);
        Y[1][b] = T(0.5) * (oy ? h0 : h1) + T(0.5) * h1;
        Y[2][b] = T(0.5) * h1 + T(0.5) * (oy ? h1 : h2);
    }
}

template <class T>
__device__ __forceinline__ void interp3x3(const Grid2D& gc, const T* yc, T my, const NodeNbr& s, T (&Y)[3][3]) {
    interp3x3_v(gc, [&](int cy, int cx) { return yc[cy * gc.nnx + cx]; }, my, s, Y);
}

}  // namespace fcfd
