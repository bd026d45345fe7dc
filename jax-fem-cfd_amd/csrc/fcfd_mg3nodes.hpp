// Node-grid transfers of the box-grid V-cycle, shared by the hex8 and hex27 hierarchies (a hex27
// node grid of nex elements is the hex8 node grid of 2 nex): trilinear prolongation P, applied
// along x, then y, then z, and its transpose R = P^T (unscaled, the 2D R = I^T convention).
#pragma once
#include "fcfd_internal.hpp"
#include "fcfd_hex.hpp"

namespace fcfd {

// coarse[I,J,K] = (R (f - mean f))[I,J,K]: along x, then y, then z; missing fine nodes are 0
struct Restrict3 {
    Grid3D gf, gc; const double* f; double* rc; const DevState* st; int l; int neumann;
    __device__ void operator()(int nc, double (&acc)[1]) const {
        const int K = nc / gc.nxy, rem = nc - K * gc.nxy;
        const int J = rem / gc.nnx, I = rem - J * gc.nnx;
        const double mf = neumann ? st->mg[l][MG_F] : 0.0;
        auto val = [&](int di, int dj, int dk) -> double {
            const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
            const bool ok = i >= 0 && j >= 0 && k >= 0 && i < gf.nnx && j < gf.nny && k < gf.nnz;
            const int id = ok ? k * gf.nxy + j * gf.nnx + i : 0;
            const double v = f[id] - mf;
            return ok ? v : 0.0;
        };
        double Y[3];
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk) {
            double Xr[3];
#pragma unroll
            for (int dj = -1; dj <= 1; ++dj) Xr[dj + 1] = val(0, dj, dk) + 0.5 * (val(-1, dj, dk) + val(1, dj, dk));
            Y[dk + 1] = Xr[1] + 0.5 * (Xr[0] + Xr[2]);
        }
        const double v = Y[1] + 0.5 * (Y[0] + Y[2]);
        rc[nc] = v;
        acc[0] += v;
    }
};

// ------------------------------------------------------------------------------ up leg
// trilinear interpolation of (yc - my) at fine node (i, j, k): x, then y, then z; at even fine
// indices both coarse neighbours coincide and 0.5 a + 0.5 a = a exactly
__device__ __forceinline__ double interp3_at(const Grid3D& gf, const Grid3D& gc, const double* yc, double my,
                                             int n) {
    const int k = n / gf.nxy, rem = n - k * gf.nxy;
    const int j = rem / gf.nnx, i = rem - j * gf.nnx;
    const int i0 = i >> 1, i1 = (i + 1) >> 1, j0 = j >> 1, j1 = (j + 1) >> 1, k0 = k >> 1, k1 = (k + 1) >> 1;
    auto at = [&](int a, int b, int c) { return yc[c * gc.nxy + b * gc.nnx + a] - my; };
    auto row = [&](int b, int c) { return 0.5 * at(i0, b, c) + 0.5 * at(i1, b, c); };
    auto lay = [&](int c) { return 0.5 * row(j0, c) + 0.5 * row(j1, c); };
    return 0.5 * lay(k0) + 0.5 * lay(k1);
}

// the interpolated values at all 27 stencil points of a node from one 3x3x3 patch of coarse
// values (fine f - 1 .. f + 1 touch coarse floor((f - 1) / 2) + {0, 1, 2}); the pairing depends
// on the parity of (i, j, k) only, with the operands and order of interp3_at
__device__ __forceinline__ void interp27(const Grid3D& gc, const double* yc, int i, int j, int k,
                                         double (&Y)[27]) {
    const int bx = (i - 1) >> 1, by = (j - 1) >> 1, bz = (k - 1) >> 1;
    int cx[3], cy[3], cz[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int x = bx + t, y = by + t, z = bz + t;
        cx[t] = x < 0 ? 0 : (x >= gc.nnx ? gc.nnx - 1 : x);
        cy[t] = y < 0 ? 0 : (y >= gc.nny ? gc.nny - 1 : y);
        cz[t] = z < 0 ? 0 : (z >= gc.nnz ? gc.nnz - 1 : z);
    }
    const bool ox = i & 1, oy = j & 1, oz = k & 1;
    // pairs (first, second) of patch indices for fine offsets -1, 0, +1:
    //   even index: (0,1) (1,1) (1,2)      odd index: (0,0) (0,1) (1,1)
    double Hy[3][3][3];  // [z patch][y offset][x offset]
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double Hx[3][3];  // [y patch][x offset]
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double* rowp = yc + (size_t)cz[c] * gc.nxy + (size_t)cy[b] * gc.nnx;
            const double p0 = rowp[cx[0]], p1 = rowp[cx[1]], p2 = rowp[cx[2]];
            Hx[b][0] = 0.5 * p0 + 0.5 * (ox ? p0 : p1);
            Hx[b][1] = 0.5 * (ox ? p0 : p1) + 0.5 * p1;
            Hx[b][2] = 0.5 * p1 + 0.5 * (ox ? p1 : p2);
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double h0 = Hx[0][a], h1 = Hx[1][a], h2 = Hx[2][a];
            Hy[c][0][a] = 0.5 * h0 + 0.5 * (oy ? h0 : h1);
            Hy[c][1][a] = 0.5 * (oy ? h0 : h1) + 0.5 * h1;
            Hy[c][2][a] = 0.5 * h1 + 0.5 * (oy ? h1 : h2);
        }
    }
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double h0 = Hy[0][b][a], h1 = Hy[1][b][a], h2 = Hy[2][b][a];
            Y[(0 * 3 + b) * 3 + a] = 0.5 * h0 + 0.5 * (oz ? h0 : h1);
            Y[(1 * 3 + b) * 3 + a] = 0.5 * (oz ? h0 : h1) + 0.5 * h1;
            Y[(2 * 3 + b) * 3 + a] = 0.5 * h1 + 0.5 * (oz ? h1 : h2);
        }
}

// p = P (yc - mean yc) (Neumann form; pinned: p = P yc), sum -> mean p
struct Prolong3 {
    Grid3D gf, gc; const double* yc; double* p; const DevState* st; int l; int neumann;
    __device__ void operator()(int n, double (&acc)[1]) const {
        const double v = interp3_at(gf, gc, yc, neumann ? st->mg[l + 1][MG_Y] : 0.0, n);
        p[n] = v;
        acc[0] += v;
    }
};

}  // namespace fcfd
