// Triquadratic (hex27) four-step timestep on structured box grids.
//
// The reference has no 3D quadratic element (simulation.py:73-74: "3D quadratic shape functions
// not yet implemented").  This path runs the chain of jax_fem_cfd/solvers/standard.py:148-165 with
// ndim = 3 on hex27 box meshes, built the way the quad9 path extends quad4 in 2D:
//   - the context's grid is its NODE grid: Grid3D with nnx = 2 nex + 1 (likewise y, z) and the
//     element counts in nex, ney, nez; element (ie, je, ke) -> (ke*ney + je)*nex + ie owns the
//     nodes (2 ie + a, 2 je + b, 2 ke + c), local index (c*3 + b)*3 + a;
//   - 3x3x3 Gauss, 27 x 27 element matrices (Ce row-blocked, FP64 or FP32);
//   - BiCGSTAB momentum (3 components), PCG pressure with none / jacobi / approx_inv or the
//     box-grid V-cycle on hex27 levels (the node-grid transfers of the hex8 hierarchy);
//   - the boundary integral on quad9 faces, mapped from the (z, y) coordinates of their nodes as
//     the reference's 3D face integral does (element_calc.py:137-163), so x-faces carry it;
//   - the mass operator is the per-component product (the reference's 3D consistent_mass_mult
//     cannot run: element_mult.py:26).
// Kernels are owner-computes node gathers: node (i, j, k) has 1, 2, 4 or 8 incident elements (by
// the parity of i, j, k); it sums, in ascending element order, its row of each, applied to that
// element's 27 nodal values read directly.
#include <algorithm>
#include <cmath>
#include "fcfd_bicgstab.hpp"
#include "fcfd_cg.hpp"
#include "fcfd_mgnodes.hpp"
#include "fcfd_mg3nodes.hpp"

namespace fcfd {

// ------------------------------------------------------------------------------ element data
__host__ __device__ constexpr int h27x(int b) { return b % 3; }
__host__ __device__ constexpr int h27y(int b) { return (b / 3) % 3; }
__host__ __device__ constexpr int h27z(int b) { return b / 9; }

// Row-blocked 27 x 27 storage with rows padded to 28 entries: entry (a, b) of element e at
// ((a * ne + e) * 28 + b).  Every row starts on a 16-byte boundary, so a node reads its row of an
// element with 7 float4 (FP32) or 14 double2 (FP64) loads instead of 27 scalar ones.
constexpr int H27_ROW = 28;
__host__ __device__ __forceinline__ size_t h27_idx(int k, int e, int ne) {
    return ((size_t)(k / 27) * ne + e) * H27_ROW + k % 27;
}
struct Ce27F64 {
    double* c; int ne;
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[27]) const {
        const double2* p = reinterpret_cast<const double2*>(c + h27_idx(a * 27, e, ne));
#pragma unroll
        for (int q = 0; q < 13; ++q) {
            const double2 t = p[q];
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
        v[26] = reinterpret_cast<const double*>(p)[26];
    }
    __device__ __forceinline__ void st(int k, int e, double v) const { c[h27_idx(k, e, ne)] = v; }
};
struct Ce27F32 {
    float* c; int ne;
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[27]) const {
        const float4* p = reinterpret_cast<const float4*>(c + h27_idx(a * 27, e, ne));
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const float4 t = p[q];
            v[4 * q] = (double)t.x;
            v[4 * q + 1] = (double)t.y;
            v[4 * q + 2] = (double)t.z;
            v[4 * q + 3] = (double)t.w;
        }
        const float4 t = p[6];
        v[24] = (double)t.x;
        v[25] = (double)t.y;
        v[26] = (double)t.z;
    }
    __device__ __forceinline__ void st(int k, int e, double v) const { c[h27_idx(k, e, ne)] = (float)v; }
};
struct Ce27Api {
    double* c;
    __device__ __forceinline__ void st(int k, int e, double v) const { c[(size_t)e * 729 + k] = v; }
};

static Grid3D make_grid27(int nex, int ney, int nez) {
    Grid3D g;
    g.nex = nex; g.ney = ney; g.nez = nez;
    g.nnx = 2 * nex + 1; g.nny = 2 * ney + 1; g.nnz = 2 * nez + 1;
    g.nxy = g.nnx * g.nny;
    g.nn = g.nxy * g.nnz;
    g.ne = nex * ney * nez;
    return g;
}

// hex27 3x3x3 Gauss data of a box element of size hx x hy x hz and its constants
// (element_calc.py:9-40 on the triquadratic tensor-product basis)
static void hex27_constants(double hx, double hy, double hz, ElemConstH27& ec) {
    const double g = std::sqrt(3.0 / 5.0);
    const double x1[3] = {-g, 0.0, g}, w1[3] = {5.0 / 9, 8.0 / 9, 5.0 / 9};
    auto L = [](double t, int a) { return a == 0 ? t * (t - 1) / 2 : (a == 1 ? 1 - t * t : t * (t + 1) / 2); };
    auto dL = [](double t, int a) { return a == 0 ? (2 * t - 1) / 2 : (a == 1 ? -2 * t : (2 * t + 1) / 2); };
    for (int q = 0; q < 27; ++q) {
        const double xi = x1[h27x(q)], eta = x1[h27y(q)], zeta = x1[h27z(q)];
        double dxi[3][27];
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int b = 0; b < 27; ++b) {
            const int a = h27x(b), by = h27y(b), c = h27z(b);
            ec.N[q * 27 + b] = L(xi, a) * L(eta, by) * L(zeta, c);
            dxi[0][b] = dL(xi, a) * L(eta, by) * L(zeta, c);
            dxi[1][b] = L(xi, a) * dL(eta, by) * L(zeta, c);
            dxi[2][b] = L(xi, a) * L(eta, by) * dL(zeta, c);
            const double X[3] = {a * hx / 2, by * hy / 2, c * hz / 2};
            for (int r = 0; r < 3; ++r)
                for (int col = 0; col < 3; ++col) J[r][col] += dxi[r][b] * X[col];
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        double Ji[3][3];
        Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        ec.wq[q] = w1[h27x(q)] * w1[h27y(q)] * w1[h27z(q)] * det;
        for (int d = 0; d < 3; ++d)
            for (int b = 0; b < 27; ++b)
                ec.dN[d][q * 27 + b] = Ji[d][0] * dxi[0][b] + Ji[d][1] * dxi[1][b] + Ji[d][2] * dxi[2][b];
    }
    for (int a = 0; a < 27; ++a) {
        for (int b = 0; b < 27; ++b) {
            double m = 0.0, l = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
            for (int q = 0; q < 27; ++q) {
                const double w = ec.wq[q], Na = ec.N[q * 27 + a], Nb = ec.N[q * 27 + b];
                m += w * Na * Nb;
                l += w * (ec.dN[0][q * 27 + a] * ec.dN[0][q * 27 + b] + ec.dN[1][q * 27 + a] * ec.dN[1][q * 27 + b] +
                          ec.dN[2][q * 27 + a] * ec.dN[2][q * 27 + b]);
                g0 += w * Na * ec.dN[0][q * 27 + b];
                g1 += w * Na * ec.dN[1][q * 27 + b];
                g2 += w * Na * ec.dN[2][q * 27 + b];
            }
            ec.Me[a * 27 + b] = m;
            ec.Le[a * 27 + b] = l;
            ec.Leoff[a * 27 + b] = a == b ? 0.0 : l;
            ec.G[0][a * 27 + b] = g0;
            ec.G[1][a * 27 + b] = g1;
            ec.G[2][a * 27 + b] = g2;
        }
        for (int q = 0; q < 27; ++q) ec.WN[a * 27 + q] = ec.wq[q] * ec.N[q * 27 + a];
    }
}

// ------------------------------------------------------------------------------ incidence
// Per axis (i on nnx): odd i -> element (i-1)/2 at local 1 only; even i -> element i/2 - 1 at
// local 2 and element i/2 at local 0.  Candidate q = qx + 2 qy + 4 qz: ascending element ids.
// The candidates are formed arithmetically inside a rolled loop (no private array is indexed at
// run time and no more than one element's row and values are live), and missing ones skipped.
struct Inc27 {
    int i, j, k;
};

__device__ __forceinline__ void node_inc_h27(const Grid3D& g, int n, Inc27& s) {
    s.k = n / g.nxy;
    const int r = n - s.k * g.nxy;
    s.j = r / g.nnx;
    s.i = r - s.j * g.nnx;
}

__device__ __forceinline__ void cand_axis(int i, int ne, int q, int& el, int& loc, bool& v) {
    const bool o = i & 1;
    el = q ? (i >> 1) : (o ? (i - 1) >> 1 : (i >> 1) - 1);
    loc = q ? 0 : (o ? 1 : 2);
    v = q ? (!o && (i >> 1) < ne) : (el >= 0);
}

// Where element (ex, ey, ez)'s local node b is read: base(ex, ey, ez) + off(b).  GIdx27: global
// node ids; CIdx27: ids in a staged tile window (k_ctile27, k_btile27); RIdx27: the element
// products already formed (k_btile27).
struct GIdx27 {
    static constexpr bool kClass = false, kRes = false;
    int nnx, nxy;
    __device__ __forceinline__ int base(int ex, int ey, int ez) const { return 2 * ez * nxy + 2 * ey * nnx + 2 * ex; }
    __device__ __forceinline__ int off(int b) const { return h27z(b) * nxy + h27y(b) * nnx + h27x(b); }
};

// Class-uniform staged tiles (k_ctile27): a block owns C27X x C27Y x C27Z nodes with an even
// origin and one wave per parity class of (i, j, k): the 64 nodes of a class have the same
// incident-element count and local rows, so the candidate loop does not diverge and the rows of
// the shared element matrices are uniform loads.  The window's x direction is stored even
// entries first, then odd ones, so the lanes of a wave (consecutive elements along x) read
// consecutive LDS words.
constexpr int C27X = 16, C27Y = 8, C27Z = 4, C27T = 512;
constexpr int V27X = C27X + 3, V27Y = C27Y + 3, V27Z = C27Z + 3, V27N = V27X * V27Y * V27Z;
constexpr int V27E = (V27X + 1) / 2;  // even-x entries of a window row
constexpr int V27IT = (V27N + C27T - 1) / C27T, V27P = V27IT * C27T;
static_assert(C27X * C27Y * C27Z == C27T && C27T == 8 * 64 && C27X == 16 && C27Y == 8 && C27Z == 4,
              "eight classes of 8 x 4 x 2 nodes, one wave each");
struct CIdx27 {
    static constexpr bool kClass = true, kRes = false;
    int wi0, wj0, wk0;  // window origin (even)
    int px, py, pz;     // the parities of the wave's nodes (uniform)
    __device__ __forceinline__ int base(int ex, int ey, int ez) const {
        return ((2 * ez - wk0) * V27Y + (2 * ey - wj0)) * V27X + (ex - (wi0 >> 1));
    }
    __device__ __forceinline__ int off(int b) const {
        return (h27z(b) * V27Y + h27y(b)) * V27X + (h27x(b) == 1 ? V27E : (h27x(b) >> 1));
    }
    __device__ __forceinline__ static int pos(int wx, int wy, int wz) {
        return (wz * V27Y + wy) * V27X + ((wx & 1) ? V27E + (wx >> 1) : (wx >> 1));
    }
    __device__ __forceinline__ int at(int i, int j, int k) const { return pos(i - wi0, j - wj0, k - wk0); }
};

// Balanced class tiles (k_btile27): the 27 (class, candidate) pairs of a 16 x 8 x 4 tile -- class
// c = px + 2 py + 4 pz has (2 - px)(2 - py)(2 - pz) incident-element candidates -- in the order of
// the candidate loop, classes ascending.  A wave forms the element products of pairs w, w + 8,
// w + 16, w + 24 for the 64 nodes of the pair's class; then the wave of class c sums its pairs in
// order.  RIdx27 hands those sums to the bodies.
struct Combo27 {
    int8_t cls[27], qx[27], qy[27], qz[27];
    int8_t m0[8], nm[8];
};
__host__ __device__ constexpr Combo27 make_combo27() {
    Combo27 t{};
    int m = 0;
    for (int c = 0; c < 8; ++c) {
        const int px = c & 1, py = (c >> 1) & 1, pz = c >> 2;
        t.m0[c] = (int8_t)m;
        for (int qz = 0; qz < 2 - pz; ++qz)
            for (int qy = 0; qy < 2 - py; ++qy)
                for (int qx = 0; qx < 2 - px; ++qx) {
                    t.cls[m] = (int8_t)c;
                    t.qx[m] = (int8_t)qx;
                    t.qy[m] = (int8_t)qy;
                    t.qz[m] = (int8_t)qz;
                    ++m;
                }
        t.nm[c] = (int8_t)(m - t.m0[c]);
    }
    return t;
}
__device__ constexpr Combo27 kCombo27 = make_combo27();
struct RIdx27 {
    static constexpr bool kClass = false, kRes = true;
    CIdx27 w;             // the window of the block (own values)
    const double* res;    // res[(c * 27 + m) * 64 + lane]: the element products of pair m, component c
    int m0, nm, lane;     // the pairs of the node's class
    __device__ __forceinline__ int at(int i, int j, int k) const { return w.at(i, j, k); }
};

// sum_q has_q sum_b A(row_q, b, e_q) X(node_b of e_q) for NC components; A from rowf(a, e, cf),
// X from val(c, ix.base(e) + ix.off(b))
template <int NC, class RowF, class ValF, class Idx>
__device__ __forceinline__ void h27_apply_ix(const Grid3D& g, const Inc27& s, RowF rowf, ValF val, double (&y)[NC],
                                             const Idx& ix) {
#pragma unroll
    for (int c = 0; c < NC; ++c) y[c] = 0.0;
    if constexpr (Idx::kRes) {
        for (int m = ix.m0; m < ix.m0 + ix.nm; ++m)
#pragma unroll
            for (int c = 0; c < NC; ++c) y[c] += ix.res[(c * 27 + m) * 64 + ix.lane];
    } else if constexpr (Idx::kClass) {
        // the candidates of the full loop below that exist for the wave's parities, in its order
        // (an odd index has only q = 0 along its axis); counts and local rows are uniform
        const int nqx = ix.px ? 1 : 2, nqy = ix.py ? 1 : 2, nqz = ix.pz ? 1 : 2;
#pragma unroll 1
        for (int qz = 0; qz < nqz; ++qz)
#pragma unroll 1
            for (int qy = 0; qy < nqy; ++qy)
#pragma unroll 1
                for (int qx = 0; qx < nqx; ++qx) {
                    int ex, cx, ey, cy, ez, cz;
                    bool vx, vy, vz;
                    cand_axis(s.i, g.nex, qx, ex, cx, vx);
                    cand_axis(s.j, g.ney, qy, ey, cy, vy);
                    cand_axis(s.k, g.nez, qz, ez, cz, vz);
                    const int a = ((ix.pz ? 1 : (qz ? 0 : 2)) * 3 + (ix.py ? 1 : (qy ? 0 : 2))) * 3 +
                                  (ix.px ? 1 : (qx ? 0 : 2));
                    const bool v = vx && vy && vz;
                    const int e = v ? (ez * g.ney + ey) * g.nex + ex : 0;
                    const int base = v ? ix.base(ex, ey, ez) : 0;
                    double cf[27];
                    rowf(a, e, cf);
                    double t[NC];
#pragma unroll
                    for (int c = 0; c < NC; ++c) t[c] = 0.0;
#pragma unroll
                    for (int b = 0; b < 27; ++b) {
                        const int id = base + ix.off(b);
#pragma unroll
                        for (int c = 0; c < NC; ++c) t[c] += cf[b] * val(c, id);
                    }
#pragma unroll
                    for (int c = 0; c < NC; ++c) y[c] += v ? t[c] : 0.0;
                }
    } else {
#pragma unroll 1
        for (int q = 0; q < 8; ++q) {
            int ex, cx, ey, cy, ez, cz;
            bool vx, vy, vz;
            cand_axis(s.i, g.nex, q & 1, ex, cx, vx);
            cand_axis(s.j, g.ney, (q >> 1) & 1, ey, cy, vy);
            cand_axis(s.k, g.nez, q >> 2, ez, cz, vz);
            if (vx && vy && vz) {
                const int e = (ez * g.ney + ey) * g.nex + ex;
                const int base = ix.base(ex, ey, ez);
                double cf[27];
                rowf((cz * 3 + cy) * 3 + cx, e, cf);
                double t[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) t[c] = 0.0;
#pragma unroll
                for (int b = 0; b < 27; ++b) {
                    const int id = base + ix.off(b);
#pragma unroll
                    for (int c = 0; c < NC; ++c) t[c] += cf[b] * val(c, id);
                }
#pragma unroll
                for (int c = 0; c < NC; ++c) y[c] += t[c];
            }
        }
    }
}

template <int NC, class RowF, class ValF>
__device__ __forceinline__ void h27_apply(const Grid3D& g, const Inc27& s, RowF rowf, ValF val, double (&y)[NC]) {
    h27_apply_ix<NC>(g, s, rowf, val, y, GIdx27{g.nnx, g.nxy});
}

// one shared 27 x 27 matrix (TR: its transpose)
template <bool TR = false, class ValF, class Idx>
__device__ __forceinline__ double h27_const_ix(const Grid3D& g, const Inc27& s, const double* __restrict__ A, ValF val,
                                               const Idx& ix) {
    double y[1];
    h27_apply_ix<1>(g, s,
                    [&](int a, int, double (&cf)[27]) {
#pragma unroll
                        for (int b = 0; b < 27; ++b) cf[b] = TR ? A[b * 27 + a] : A[a * 27 + b];
                    },
                    [&](int, int id) { return val(id); }, y, ix);
    return y[0];
}
template <bool TR = false, class ValF>
__device__ __forceinline__ double h27_const(const Grid3D& g, const Inc27& s, const double* __restrict__ A, ValF val) {
    return h27_const_ix<TR>(g, s, A, val, GIdx27{g.nnx, g.nxy});
}

// The class-uniform staged tile kernel (CIdx27), 512 threads.  Body::stage(id, v) forms the
// Body::NV values of node id that the element products read (e.g. p = r + beta (p_old - omega
// q_old)), once per window node; every load of the block is issued before the first LDS store,
// and window nodes outside the grid hold 0 (no existing element reads them).  The body then runs
// per owned node with the window in LDS, in the operand order of a direct gather (bitwise its
// results).  Wave w = the nodes (i0 + 2a + px, j0 + 2b + py, k0 + 2d + pz) of class
// w = px + 2 py + 4 pz, lane = a + 8 b + 32 d.
template <int NR, class Body>
__global__ __launch_bounds__(C27T) void k_ctile27(Grid3D g, int ntx, int nty, const int* __restrict__ active,
                                                  Body body, double* __restrict__ partials, int pstride) {
    constexpr int NV = Body::NV;
    if (active && *active == 0) return;
    __shared__ double sm[NV][V27P];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int tz = lb / (ntx * nty), rem = lb - tz * (ntx * nty), ty = rem / ntx, tx = rem - ty * ntx;
    const int i0 = tx * C27X, j0 = ty * C27Y, k0 = tz * C27Z;
    const int cls = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const CIdx27 ix{i0 - 2, j0 - 2, k0 - 2, cls & 1, (cls >> 1) & 1, cls >> 2};
    double v[V27IT][NV];
    bool in[V27IT];
    int at[V27IT];
#pragma unroll
    for (int it = 0; it < V27IT; ++it) {
        const int w = threadIdx.x + it * C27T;
        const int wz = w / (V27X * V27Y), wr = w - wz * (V27X * V27Y), wy = wr / V27X, wx = wr - wy * V27X;
        const int i = ix.wi0 + wx, j = ix.wj0 + wy, k = ix.wk0 + wz;
        const bool ok = w < V27N && i >= 0 && i < g.nnx && j >= 0 && j < g.nny && k >= 0 && k < g.nnz;
        in[it] = ok;
        at[it] = w < V27N ? CIdx27::pos(wx, wy, wz) : w;
        body.stage(ok ? (k * g.nny + j) * g.nnx + i : 0, v[it]);
    }
#pragma unroll
    for (int it = 0; it < V27IT; ++it)
#pragma unroll
        for (int c = 0; c < NV; ++c) sm[c][at[it]] = in[it] ? v[it][c] : 0.0;
    __syncthreads();
    double acc[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < (NR > 0 ? NR : 1); ++r) acc[r] = 0.0;
    const int lane = threadIdx.x & 63;
    const int i = i0 + 2 * (lane & 7) + ix.px, j = j0 + 2 * ((lane >> 3) & 3) + ix.py,
              k = k0 + 2 * (lane >> 5) + ix.pz;
    if (i < g.nnx && j < g.nny && k < g.nnz) body((k * g.nny + j) * g.nnx + i, Inc27{i, j, k}, ix, sm, acc);
    if constexpr (NR > 0) reduce_epilogue<NR>(acc, partials, pstride, lb);
}

template <int NR, class Body, class Fin>
static void launch_ctile27(Ctx& c, const Grid3D& g, const int* active, const Body& body, const Fin& fin) {
    const int ntx = (g.nnx + C27X - 1) / C27X, nty = (g.nny + C27Y - 1) / C27Y, ntz = (g.nnz + C27Z - 1) / C27Z;
    const int nb = ntx * nty * ntz;
    if (NR > 0 && nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
    k_ctile27<NR><<<nb, C27T, 0, c.stream>>>(g, ntx, nty, active, body, c.partials, c.pstride);
    if constexpr (NR > 0)
        finalize_kernel<NR><<<1, FIN_BLOCK, 0, c.stream>>>(nb, fin_gate(active, fin, 0), fin, c.partials, c.pstride);
}

// The balanced class tile kernel: k_ctile27's tile, window and staging; the element products of
// the tile's 27 (class, candidate) pairs spread over the 8 waves (3-4 pairs each instead of 1-8),
// kept in LDS and summed per node in candidate order (RIdx27), so bitwise the k_ctile27 results.
// Bodies with one element product of NC components read from the first NC staged vectors.
template <int NR, class Body>
__global__ __launch_bounds__(C27T) void k_btile27(Grid3D g, int ntx, int nty, const int* __restrict__ active,
                                                  Body body, double* __restrict__ partials, int pstride) {
    constexpr int NV = Body::NV, NC = Body::NC;
    static_assert(NC <= NV, "the products read the first NC staged vectors");
    if (active && *active == 0) return;
    __shared__ double sm[NV][V27P];
    __shared__ double res[NC * 27 * 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int tz = lb / (ntx * nty), rem = lb - tz * (ntx * nty), ty = rem / ntx, tx = rem - ty * ntx;
    const int i0 = tx * C27X, j0 = ty * C27Y, k0 = tz * C27Z;
    const CIdx27 ix{i0 - 2, j0 - 2, k0 - 2, 0, 0, 0};
    double v[V27IT][NV];
    bool in[V27IT];
    int at[V27IT];
#pragma unroll
    for (int it = 0; it < V27IT; ++it) {
        const int w = threadIdx.x + it * C27T;
        const int wz = w / (V27X * V27Y), wr = w - wz * (V27X * V27Y), wy = wr / V27X, wx = wr - wy * V27X;
        const int i = ix.wi0 + wx, j = ix.wj0 + wy, k = ix.wk0 + wz;
        const bool ok = w < V27N && i >= 0 && i < g.nnx && j >= 0 && j < g.nny && k >= 0 && k < g.nnz;
        in[it] = ok;
        at[it] = w < V27N ? CIdx27::pos(wx, wy, wz) : w;
        body.stage(ok ? (k * g.nny + j) * g.nnx + i : 0, v[it]);
    }
#pragma unroll
    for (int it = 0; it < V27IT; ++it)
#pragma unroll
        for (int c = 0; c < NV; ++c) sm[c][at[it]] = in[it] ? v[it][c] : 0.0;
    __syncthreads();
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int la = lane & 7, lm = (lane >> 3) & 3, lh = lane >> 5;
    for (int m = wv; m < 27; m += 8) {
        const int c = kCombo27.cls[m];
        const int px = c & 1, py = (c >> 1) & 1, pz = c >> 2;
        const int qx = kCombo27.qx[m], qy = kCombo27.qy[m], qz = kCombo27.qz[m];
        const int i = i0 + 2 * la + px, j = j0 + 2 * lm + py, k = k0 + 2 * lh + pz;
        int ex, cx, ey, cy, ez, cz;
        bool vx, vy, vz;
        cand_axis(i, g.nex, qx, ex, cx, vx);
        cand_axis(j, g.ney, qy, ey, cy, vy);
        cand_axis(k, g.nez, qz, ez, cz, vz);
        const int a = ((pz ? 1 : (qz ? 0 : 2)) * 3 + (py ? 1 : (qy ? 0 : 2))) * 3 + (px ? 1 : (qx ? 0 : 2));
        // cand_axis assumes a node of the grid: a lane past the grid's far edge (ragged last tile)
        // would get element index ne along that axis, so such lanes are masked here
        const bool ok = vx && vy && vz && i < g.nnx && j < g.nny && k < g.nnz;
        const int e = ok ? (ez * g.ney + ey) * g.nex + ex : 0;
        const int base = ok ? ix.base(ex, ey, ez) : 0;
        double cf[27];
        body.row(a, e, cf);
        double t[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) t[q] = 0.0;
#pragma unroll
        for (int b = 0; b < 27; ++b) {
            const int id = base + ix.off(b);
#pragma unroll
            for (int q = 0; q < NC; ++q) t[q] += cf[b] * sm[q][id];
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) res[(q * 27 + m) * 64 + lane] = ok ? t[q] : 0.0;
    }
    __syncthreads();
    double acc[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < (NR > 0 ? NR : 1); ++r) acc[r] = 0.0;
    {
        const int px = wv & 1, py = (wv >> 1) & 1, pz = wv >> 2;
        const int i = i0 + 2 * la + px, j = j0 + 2 * lm + py, k = k0 + 2 * lh + pz;
        const RIdx27 rx{ix, res, kCombo27.m0[wv], kCombo27.nm[wv], lane};
        if (i < g.nnx && j < g.nny && k < g.nnz) body((k * g.nny + j) * g.nnx + i, Inc27{i, j, k}, rx, sm, acc);
    }
    if constexpr (NR > 0) reduce_epilogue<NR>(acc, partials, pstride, lb);
}

template <int NR, class Body, class Fin>
static void launch_btile27(Ctx& c, const Grid3D& g, const int* active, const Body& body, const Fin& fin) {
    const int ntx = (g.nnx + C27X - 1) / C27X, nty = (g.nny + C27Y - 1) / C27Y, ntz = (g.nnz + C27Z - 1) / C27Z;
    const int nb = ntx * nty * ntz;
    if (NR > 0 && nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
    k_btile27<NR><<<nb, C27T, 0, c.stream>>>(g, ntx, nty, active, body, c.partials, c.pstride);
    if constexpr (NR > 0)
        finalize_kernel<NR><<<1, FIN_BLOCK, 0, c.stream>>>(nb, fin_gate(active, fin, 0), fin, c.partials, c.pstride);
}


// ------------------------------------------------------------------------------ Ce
// Ce[a][b] = sum_q w_q N_qa (U_q . grad N_qb + beta div U_q N_qb)   (element_calc.py:165-185)
// A block builds CE27_E consecutive elements with 27 threads each: thread (el, b) stages the
// element's nodal U, thread (el, q) forms U, V, W and div U at Gauss point q (sums over b in order),
// then thread (el, b) forms the advection term of column b at the 27 points and the 27 entries of
// column b.  For each row a the block's stores cover CE27_E * 27 consecutive values of the
// row-blocked layout, so they coalesce.
constexpr int CE27_E = 9;  // 243 of 256 threads
template <class S>
__global__ __launch_bounds__(256) void k_build_ce_h27(Grid3D g, const ElemConstH27* __restrict__ ec,
                                                     const double* __restrict__ U, double beta, S store) {
    __shared__ double un[CE27_E][3][27];
    __shared__ double qv[CE27_E][4][27];
    const int el = threadIdx.x / 27, b = threadIdx.x - el * 27;
    const int e = xcd_block(blockIdx.x, gridDim.x) * CE27_E + el;
    const bool live = el < CE27_E && e < g.ne;
    if (live) {
        const int ke = e / (g.nex * g.ney), r = e - ke * g.nex * g.ney;
        const int je = r / g.nex, ie = r - je * g.nex;
        const int id = 2 * ke * g.nxy + 2 * je * g.nnx + 2 * ie + h27z(b) * g.nxy + h27y(b) * g.nnx + h27x(b);
        un[el][0][b] = U[id];
        un[el][1][b] = U[id + g.nn];
        un[el][2][b] = U[id + 2 * (size_t)g.nn];
    }
    __syncthreads();
    if (live) {
        const int q = b;
        double uq = 0.0, vq = 0.0, wq = 0.0, dq = 0.0;
#pragma unroll
        for (int bb = 0; bb < 27; ++bb) {
            const double u = un[el][0][bb], v = un[el][1][bb], w = un[el][2][bb];
            const double Nb = ec->N[q * 27 + bb];
            uq += Nb * u;
            vq += Nb * v;
            wq += Nb * w;
            dq += ec->dN[0][q * 27 + bb] * u + ec->dN[1][q * 27 + bb] * v + ec->dN[2][q * 27 + bb] * w;
        }
        qv[el][0][q] = uq;
        qv[el][1][q] = vq;
        qv[el][2][q] = wq;
        qv[el][3][q] = dq;
    }
    __syncthreads();
    if (!live) return;
    double adv[27];
#pragma unroll
    for (int q = 0; q < 27; ++q)
        adv[q] = qv[el][0][q] * ec->dN[0][q * 27 + b] + qv[el][1][q] * ec->dN[1][q * 27 + b] +
                 qv[el][2][q] * ec->dN[2][q * 27 + b] + beta * qv[el][3][q] * ec->N[q * 27 + b];
    for (int a = 0; a < 27; ++a) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < 27; ++q) t += ec->WN[a * 27 + q] * adv[q];
        store.st(a * 27 + b, e, t);
    }
}

template <class S>
static void launch_build_ce_h27(Ctx& c, const Grid3D& g, const double* U, double beta, S store) {
    const int nb = (g.ne + CE27_E - 1) / CE27_E;
    k_build_ce_h27<<<nb, 256, 0, c.stream>>>(g, c.eh27_dev, U, beta, store);
    FCFD_CHECK_LAUNCH();
}

static void build_ce_h27(Ctx& c, const double* U, double beta) {
    if (c.ce_format == FCFD_CE_F64) launch_build_ce_h27(c, c.g3, U, beta, Ce27F64{(double*)c.ce, c.g3.ne});
    else launch_build_ce_h27(c, c.g3, U, beta, Ce27F32{(float*)c.ce, c.g3.ne});
}

void build_ce_h27_api(Ctx& c, const double* U, double beta, double* Ce) {
    launch_build_ce_h27(c, c.g3, U, beta, Ce27Api{Ce});
}

// ------------------------------------------------------------------------------ momentum
template <class S>
struct MomCoef27 {
    Grid3D g;
    const ElemConstH27* __restrict__ ec;
    S ce;
    double cM, cL, cC;
    __device__ __forceinline__ void row(int a, int e, double (&cf)[27]) const {
        double v[27];
        ce.ld_row(a, e, v);
#pragma unroll
        for (int b = 0; b < 27; ++b) cf[b] = cM * ec->Me[a * 27 + b] + cL * ec->Le[a * 27 + b] + cC * v[b];
    }
};

// rhs = R U_n - G P_n + F (Dirichlet rows U_d), x0 = U_n with Dirichlet values, r0 = rhs - A x0,
// rhat = p = q = r0; reductions b.b, r0.r0 (the quad4 BInit on hex27 gathers).  Staged: U_n (3
// components), x0 (3) and P_n.
template <class S>
struct BInit27 {
    static constexpr int NV = 7;
    MomCoef27<S> lhs, rhs;
    const double *xn, *ud, *Pn, *F;
    const uint8_t* mask;
    double *x, *r, *rhat, *p, *q;
    __device__ void stage(int id, double (&v)[NV]) const {
        const int nn = lhs.g.nn;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = id + c * nn;
            const double xv = xn[i];
            v[c] = xv;
            v[3 + c] = (mask && mask[i]) ? ud[i] : xv;
        }
        v[6] = Pn[id];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[2]) const {
        const Grid3D& g = lhs.g;
        const int nn = g.nn;
        double b[3], ax[3];
        h27_apply_ix<3>(g, s, [&](int a, int e, double (&cf)[27]) { rhs.row(a, e, cf); },
                        [&](int c, int id) { return sm[c][id]; }, b, ix);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            b[c] -= h27_const_ix(g, s, lhs.ec->G[c], [&](int id) { return sm[6][id]; }, ix);
            if (F) b[c] += F[n + c * nn];
        }
        h27_apply_ix<3>(g, s, [&](int a, int e, double (&cf)[27]) { lhs.row(a, e, cf); },
                        [&](int c, int id) { return sm[3 + c][id]; }, ax, ix);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
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

// p = r + beta (p_old - omega q_old) (staged), q = A p, rhat . q
template <class Op>
struct BK1_27 {
    static constexpr int NV = 3, NC = 3;
    Op op;
    const uint8_t* mask;
    const double *r, *pold, *qold, *rhat;
    double *pnew, *qnew;
    const DevState* st;
    __device__ void stage(int id, double (&v)[NV]) const {
        const double beta = st->b_beta, omega = st->b_omega;
        const int nn = op.g.nn;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = id + c * nn;
            v[c] = r[i] + beta * (pold[i] - omega * qold[i]);
        }
    }
    __device__ __forceinline__ void row(int a, int e, double (&cf)[27]) const { op.row(a, e, cf); }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[1]) const {
        const Grid3D& g = op.g;
        const int nn = g.nn;
        double y[3];
        h27_apply_ix<3>(g, s, [&](int a, int e, double (&cf)[27]) { op.row(a, e, cf); },
                        [&](int c, int id) { return sm[c][id]; }, y, ix);
        const int own = ix.at(s.i, s.j, s.k);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = n + c * nn;
            const double pc = sm[c][own];
            const double yc = (mask && mask[i]) ? pc : y[c];
            pnew[i] = pc;
            qnew[i] = yc;
            acc[0] += rhat[i] * yc;
        }
    }
};

// s = r - alpha q (staged), t = A s, s.s, t.s, t.t
template <class Op>
struct BK2_27 {
    static constexpr int NV = 3, NC = 3;
    Op op;
    const uint8_t* mask;
    const double *r, *q;
    double *sv, *tv;
    const DevState* st;
    __device__ void stage(int id, double (&v)[NV]) const {
        const double alpha = st->b_alpha;
        const int nn = op.g.nn;
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = r[id + c * nn] - alpha * q[id + c * nn];
    }
    __device__ __forceinline__ void row(int a, int e, double (&cf)[27]) const { op.row(a, e, cf); }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[3]) const {
        const Grid3D& g = op.g;
        const int nn = g.nn;
        double y[3];
        h27_apply_ix<3>(g, s, [&](int a, int e, double (&cf)[27]) { op.row(a, e, cf); },
                        [&](int c, int id) { return sm[c][id]; }, y, ix);
        const int own = ix.at(s.i, s.j, s.k);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = n + c * nn;
            const double sc = sm[c][own];
            const double yc = (mask && mask[i]) ? sc : y[c];
            sv[i] = sc;
            tv[i] = yc;
            acc[0] += sc * sc;
            acc[1] += yc * sc;
            acc[2] += yc * yc;
        }
    }
};

template <class S>
static int solve_momentum_h27_t(Ctx& c, S ce, const double* U_n, const double* P_n, double rho, double mu, double dt,
                                const double* F, double* U_hat) {
    const double theta = 0.5;
    const Grid3D& g = c.g3;
    const int nn = g.nn;
    using Op = MomCoef27<S>;
    Op lhs{g, c.eh27_dev, ce, rho / dt, theta * mu, theta * rho};
    Op rhs{g, c.eh27_dev, ce, rho / dt, -(1.0 - theta) * mu, -(1.0 - theta) * rho};
    const int maxit = c.maxit_mom > 0 ? c.maxit_mom : 10 * 3 * nn;
    launch_ctile27<2>(c, g, nullptr,
                      BInit27<S>{lhs, rhs, U_n, c.ud_full, P_n, F, c.dmask, c.m_x, c.m_r, c.m_rhat, c.m_p[0], c.m_q[0]},
                      BInitFin{c.st, c.tol, maxit});
    FCFD_CHECK_LAUNCH();
    const int* act = &c.st->b_active;
    auto pair = [&]() {
        for (int o = 0; o < 2; ++o) {
            const int nw = o ^ 1;
            launch_btile27<1>(c, g, act,
                             BK1_27<Op>{lhs, c.dmask, c.m_r, c.m_p[o], c.m_q[o], c.m_rhat, c.m_p[nw], c.m_q[nw], c.st},
                             BK1Fin{c.st});
            launch_btile27<3>(c, g, act, BK2_27<Op>{lhs, c.dmask, c.m_r, c.m_q[nw], c.m_s, c.m_t, c.st},
                             BK2Fin{c.st});
            launch_node<2>(c, nn, act, BK3<3>{nn, c.m_p[nw], c.m_s, c.m_t, c.m_rhat, c.m_x, c.m_r, c.st},
                           BK3Fin{c.st});
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
    if (h[0] > 0) c.pred_mom = h[0];
    if (U_hat != c.m_x)
        FCFD_HIP(hipMemcpyAsync(U_hat, c.m_x, sizeof(double) * 3 * nn, hipMemcpyDeviceToDevice, c.stream));
    return h[0];
}

int solve_momentum_h27(Ctx& c, const double* U_n, const double* P_n, double rho, double mu, double dt, const double* F,
                       double* U_hat) {
    build_ce_h27(c, U_n, 0.5);
    if (c.ce_format == FCFD_CE_F64)
        return solve_momentum_h27_t(c, Ce27F64{(double*)c.ce, c.g3.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
    return solve_momentum_h27_t(c, Ce27F32{(float*)c.ce, c.g3.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
}

// ------------------------------------------------------------------------------ projection
// Uout = Uin + c G P / M (c = dt/rho: U*; c = -dt/rho with Dirichlet values: the final U)
struct Proj27Body {
    static constexpr int NV = 1;
    Grid3D g;
    const ElemConstH27* ec;
    const double *Uin, *P, *M, *ud;
    const uint8_t* mask;
    double* Uout;
    double c;
    __device__ void stage(int id, double (&v)[NV]) const { v[0] = P[id]; }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&)[1]) const {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double gd = h27_const_ix(g, s, ec->G[d], [&](int id) { return sm[0][id]; }, ix);
            const size_t i = n + (size_t)d * g.nn;
            double u = Uin[i] + (c * gd) / M[i];
            if (mask && mask[i]) u = ud[i];
            Uout[i] = u;
        }
    }
};

// ------------------------------------------------------------------------------ pressure
struct Prhs27Body {
    static constexpr int NV = 3;
    Grid3D g; const ElemConstH27* ec; const double* Us; const double* Fp; const uint8_t* pin; double* rhs; double c;
    __device__ void stage(int id, double (&v)[NV]) const {
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] = Us[id + (size_t)d * g.nn];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[1]) const {
        double div = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) div += h27_const_ix<true>(g, s, ec->G[d], [&](int id) { return sm[d][id]; }, ix);
        double v = c * (div - (Fp ? Fp[n] : 0.0));
        if (pin && pin[n]) v = 0.0;
        rhs[n] = v;
        acc[0] += v;
    }
};

struct CgInit27 {
    static constexpr int NV = 1;
    Grid3D g; const double* Le; const uint8_t* pin; const double* b; const double* bmean; const double* x0in;
    int zero_pins; double* x; double* r;
    __device__ void stage(int id, double (&v)[NV]) const {
        const double xv = x0in[id];
        v[0] = (zero_pins && pin && pin[id]) ? 0.0 : xv;
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[3]) const {
        const bool pn = pin && pin[n];
        const double ax = pn ? 0.0 : h27_const_ix(g, s, Le, [&](int id) { return sm[0][id]; }, ix);
        const double bn = b[n] - (bmean ? *bmean : 0.0);
        const double rn = bn - ax;
        x[n] = sm[0][ix.at(s.i, s.j, s.k)];
        r[n] = rn;
        acc[0] += bn * bn;
        acc[1] += rn * rn;
        acc[2] += rn;
    }
};

struct CgK1_27 {
    static constexpr int NV = 1, NC = 1;
    Grid3D g; const double* Le; const uint8_t* pin; const double* z; const double* pold; double* pnew; double* Ap;
    const DevState* st;
    __device__ void stage(int id, double (&v)[NV]) const { v[0] = z[id] + st->c_beta * pold[id]; }
    __device__ __forceinline__ void row(int a, int, double (&cf)[27]) const {
#pragma unroll
        for (int b = 0; b < 27; ++b) cf[b] = Le[a * 27 + b];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[1]) const {
        const double a =
            (pin && pin[n]) ? 0.0 : h27_const_ix(g, s, Le, [&](int id) { return sm[0][id]; }, ix);
        const double pn = sm[0][ix.at(s.i, s.j, s.k)];
        pnew[n] = pn;
        Ap[n] = a;
        acc[0] += pn * a;
    }
};

// two-term Neumann-series approximate inverse (standard.py:225-234)
struct ApproxInv27a {
    static constexpr int NV = 1, NC = 1;
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; double* d1l;
    __device__ void stage(int id, double (&v)[NV]) const { v[0] = r[id] / Ld[id]; }
    __device__ __forceinline__ void row(int a, int, double (&cf)[27]) const {
#pragma unroll
        for (int b = 0; b < 27; ++b) cf[b] = Le[a * 27 + b];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&)[1]) const {
        const double v = h27_const_ix(g, s, Le, [&](int id) { return sm[0][id]; }, ix) / Ld[n];
        d1l[n] = (pin && pin[n]) ? 0.0 : v;
    }
};
struct ApproxInv27b {
    static constexpr int NV = 1, NC = 1;
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; const double* d1l; double* z;
    __device__ void stage(int id, double (&v)[NV]) const { v[0] = d1l[id]; }
    __device__ __forceinline__ void row(int a, int, double (&cf)[27]) const {
#pragma unroll
        for (int b = 0; b < 27; ++b) cf[b] = Le[a * 27 + b];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&)[1]) const {
        double l2 = h27_const_ix(g, s, Le, [&](int id) { return sm[0][id]; }, ix);
        if (pin && pin[n]) l2 = 0.0;
        const double d1x = r[n] / Ld[n];
        z[n] = (3.0 * d1x - 3.0 * d1l[n]) + l2 / Ld[n];
    }
};

// The PCG iteration's forms of the two stages (pinned form, no flexible term).  Stage 1 with
// CgK2's x / r update (jax_iterative_solvers.py:108-109) fused in: the window stages r_new /
// L_diag with r_new = r - alpha Ap formed as CgK2 forms it, and the block stores x + alpha p and
// r_new of its own nodes (into rout: rin stays intact, the neighbouring blocks stage it) with the
// partial of r_new.r_new.  The stencil is formed before any store: a store ahead of it would force
// Le's uniform loads off the scalar path (the compiler cannot rule out that x or rout alias Le).
struct ApproxInv27aK2 {
    static constexpr int NV = 1, NC = 1;
    Grid3D g; const double* Le; const double* rin; const double* Ap; const double* p; double* x; double* rout;
    const double* Ld; const uint8_t* pin; double* d1l; const DevState* st;
    __device__ void stage(int id, double (&v)[NV]) const {
        const double rk = rin[id] - st->c_alpha * Ap[id];
        v[0] = rk / Ld[id];
    }
    __device__ __forceinline__ void row(int a, int, double (&cf)[27]) const {
#pragma unroll
        for (int b = 0; b < 27; ++b) cf[b] = Le[a * 27 + b];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[1]) const {
        const double v = h27_const_ix(g, s, Le, [&](int id) { return sm[0][id]; }, ix) / Ld[n];
        const double a = st->c_alpha;
        const double xn = x[n] + a * p[n];
        const double rk = rin[n] - a * Ap[n];
        x[n] = xn;
        rout[n] = rk;
        acc[0] += rk * rk;
        d1l[n] = (pin && pin[n]) ? 0.0 : v;
    }
};
// Stage 2 with the CG update that follows it (CgK3 with no mean and no flexible term): z_new = M r
// stored, partial r.z_new
struct ApproxInv27bK3 {
    static constexpr int NV = 1, NC = 1;
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; const double* d1l; double* z;
    __device__ void stage(int id, double (&v)[NV]) const { v[0] = d1l[id]; }
    __device__ __forceinline__ void row(int a, int, double (&cf)[27]) const {
#pragma unroll
        for (int b = 0; b < 27; ++b) cf[b] = Le[a * 27 + b];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[2]) const {
        double l2 = h27_const_ix(g, s, Le, [&](int id) { return sm[0][id]; }, ix);
        if (pin && pin[n]) l2 = 0.0;
        const double d1x = r[n] / Ld[n];
        const double zk = (3.0 * d1x - 3.0 * d1l[n]) + l2 / Ld[n];
        z[n] = zk;
        acc[0] += r[n] * zk;
    }
};

// ------------------------------------------------------------------------------ multigrid
// The box-grid V-cycle (multigrid.py:231-330 carried to 3D) on hex27 levels of nex >> l elements:
// the level operators are the 27 x 27 hex27 Laplacians, the transfers the node-grid trilinear P
// and R = P^T shared with hex8 (a hex27 node grid of nex elements is a hex8 node grid of 2 nex).
// The prolongation p = P y is formed by its own kernel (into the level's f, dead by then) and the
// Jacobi reads it, since a hex27 row spans a 5x5x5 node neighbourhood.

// f = r - L x: pinned (x = 2/3 r / L_diag with the pins 0, pinned rows of L x 0) or Neumann (x
// and r mean-projected)
struct DownRes27 {
    static constexpr int NV = 1, NC = 1;
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; double* x; double* f;
    const DevState* st; int l; int neumann;
    // the pre-smoothed x (pinned: 2/3 r / L_diag, 0 at the pins) or the mean-projected x (Neumann)
    __device__ void stage(int id, double (&v)[NV]) const {
        if (neumann) {
            v[0] = x[id] - st->mg[l][MG_X];
        } else {
            const double xv = (2.0 / 3.0 * r[id]) / Ld[id];
            v[0] = (pin && pin[id]) ? 0.0 : xv;
        }
    }
    __device__ __forceinline__ void row(int a, int, double (&cf)[27]) const {
#pragma unroll
        for (int b = 0; b < 27; ++b) cf[b] = Le[a * 27 + b];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[1]) const {
        auto xv = [&](int id) { return sm[0][id]; };
        double fv;
        if (neumann) {
            const double mr = st->mg[l][MG_R];
            fv = (r[n] - mr) - h27_const_ix(g, s, Le, xv, ix);
        } else {
            x[n] = xv(ix.at(s.i, s.j, s.k));
            const bool pn = pin && pin[n];
            const double rh = pn ? 0.0 : r[n];
            const double lx = pn ? 0.0 : h27_const_ix(g, s, Le, xv, ix);
            fv = rh - lx;
        }
        f[n] = fv;
        acc[0] += fv;
    }
};

// x <- x + p (p = P y of the coarser level); one damped Jacobi (multigrid.py:214-228, :306-326)
struct Jacobi27 {
    static constexpr int NV = 1, NC = 1;
    Grid3D g; const double* Leoff; const double* Ld; const uint8_t* pin; const double* r; const double* x;
    const double* p; double* y; const DevState* st; int l; int neumann;
    // x + P y (0 at the pins), or with the means projected out (Neumann)
    __device__ void stage(int id, double (&v)[NV]) const {
        if (neumann) {
            const double mx = st->mg[l][MG_X], mp = st->mg[l][MG_P];
            v[0] = (x[id] - mx) + (p[id] - mp);
        } else {
            const double xv = x[id] + p[id];
            v[0] = (pin && pin[id]) ? 0.0 : xv;
        }
    }
    __device__ __forceinline__ void row(int a, int, double (&cf)[27]) const {
#pragma unroll
        for (int b = 0; b < 27; ++b) cf[b] = Leoff[a * 27 + b];
    }
    template <class SM, class IX>
    __device__ void operator()(int n, const Inc27& s, const IX& ix, SM sm, double (&acc)[1]) const {
        auto xh = [&](int id) { return sm[0][id]; };
        const double rh = neumann ? r[n] - (l == 0 ? 0.0 : st->mg[l][MG_R]) : r[n];
        const double off = h27_const_ix(g, s, Leoff, xh, ix);
        double v = (OMEGA * (rh - off)) / Ld[n] + (1.0 - OMEGA) * xh(ix.at(s.i, s.j, s.k));
        if (!neumann && pin && pin[n]) v = 0.0;
        y[n] = v;
        acc[0] += v;
    }
};

// Pinned coarse level: CG (M = identity) to tol, maxiter 1000, pinned rows zeroed
// (multigrid.py:278-287), one workgroup, the search direction in LDS
__global__ __launch_bounds__(1024) void k_coarse_cg27(Grid3D g, const ElemConstH27* __restrict__ ec,
                                                      const uint8_t* pin, const double* r, double* x, double* cr,
                                                      double* cAp, DevState* st, double tol, int maxit,
                                                      const int* active) {
    if (active && *active == 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* p = reinterpret_cast<double*>(smem_raw);
    const int nn = g.nn, bs = blockDim.x;
    double g2[1] = {0.0};
    for (int n = threadIdx.x; n < nn; n += bs) {
        const double b = (pin && pin[n]) ? 0.0 : r[n];
        cr[n] = b; p[n] = b; x[n] = 0.0;
        g2[0] += b * b;
    }
    block_sum<1>(g2);  // barrier: p complete
    double gamma = g2[0];
    const double atol2 = fmax(tol * tol * gamma, 0.0);
    int k = 0;
    while (gamma > atol2 && k < maxit) {
        double pap[1] = {0.0};
        for (int n = threadIdx.x; n < nn; n += bs) {
            Inc27 s;
            node_inc_h27(g, n, s);
            const double a = (pin && pin[n]) ? 0.0 : h27_const(g, s, ec->Le, [&](int m) { return p[m]; });
            cAp[n] = a;
            pap[0] += p[n] * a;
        }
        block_sum<1>(pap);  // barrier: every read of p is done
        const double alpha = gamma / pap[0];
        double rr[1] = {0.0};
        for (int n = threadIdx.x; n < nn; n += bs) {
            x[n] = x[n] + alpha * p[n];
            const double rn = cr[n] - alpha * cAp[n];
            cr[n] = rn;
            rr[0] += rn * rn;
        }
        block_sum<1>(rr);
        const double beta = rr[0] / gamma;
        for (int n = threadIdx.x; n < nn; n += bs) p[n] = cr[n] + beta * p[n];
        __syncthreads();
        gamma = rr[0];
        ++k;
    }
    for (int n = threadIdx.x; n < nn; n += bs)
        if (pin && pin[n]) x[n] = 0.0;
    if (threadIdx.x == 0) st->coarse_k = k;
}

// r0: the fine residual (its mean in mg[0][MG_R] for the Neumann form); result in lv[0].y
static const double* vcycle_h27(Ctx& c, const double* r0, const int* active) {
    const int N = c.nlev;
    const int neu = !c.set_zeroP;
    const int ld = N - 1;
    auto run = [&](int n, const auto& body, const MeanFin& fin) {
        if (neu) launch_node<1>(c, n, active, body, fin);
        else launch_node<0>(c, n, active, body, fin);
    };
    // the down leg measured faster on the balanced tiles, the up leg on the class tiles
    auto run_tile = [&](const Grid3D& g, const auto& body, const MeanFin& fin) {
        if (neu) launch_btile27<1>(c, g, active, body, fin);
        else launch_btile27<0>(c, g, active, body, fin);
    };
    auto run_ctile = [&](const Grid3D& g, const auto& body, const MeanFin& fin) {
        if (neu) launch_ctile27<1>(c, g, active, body, fin);
        else launch_ctile27<0>(c, g, active, body, fin);
    };
    auto rl = [&](int l) -> const double* { return l == 0 ? r0 : c.lv[l].r; };
    for (int l = 0; l < ld; ++l) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        const int n = L.g3.nn;
        if (neu)
            launch_node<1>(c, n, active, DownX<double, double>{rl(l), L.Ld, L.x, c.st, l},
                           MeanFin{&c.st->mg[l][MG_X], nullptr, n});
        run_tile(L.g3, DownRes27{L.g3, L.eh27_dev->Le, rl(l), L.Ld, L.pin, L.x, L.f, c.st, l, neu},
                 MeanFin{&c.st->mg[l][MG_F], nullptr, n});
        run(Lc.g3.nn, Restrict3{L.g3, Lc.g3, L.f, Lc.r, c.st, l, neu},
            MeanFin{&c.st->mg[l][MG_C], &c.st->mg[l + 1][MG_R], Lc.g3.nn});
    }
    {
        Level& C = c.lv[ld];
        if (neu) {
            k_coarse_jacobi<double, double><<<1, 1024, 0, c.stream>>>(C.g3.nn, C.Ld, rl(ld), C.y, c.st, ld, active);
        } else if (c.coarse_inv32) {
            coarse_dense64(c, C.g3.nn, C.pin, rl(ld), C.y, active);
        } else {
            const size_t shm = sizeof(double) * (size_t)C.g3.nn;
            k_coarse_cg27<<<1, 1024, shm, c.stream>>>(C.g3, C.eh27_dev, C.pin, rl(ld), C.y, c.cc_r, c.cc_Ap, c.st,
                                                      c.tol, 1000, active);
        }
        FCFD_CHECK_LAUNCH();
    }
    for (int l = ld - 1; l >= 0; --l) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        const int n = L.g3.nn;
        run(n, Prolong3{L.g3, Lc.g3, Lc.y, L.f, c.st, l, neu}, MeanFin{&c.st->mg[l][MG_P], nullptr, n});
        run_ctile(L.g3, Jacobi27{L.g3, L.eh27_dev->Leoff, L.Ld, L.pin, rl(l), L.x, L.f, L.y, c.st, l, neu},
                 MeanFin{&c.st->mg[l][MG_Y], nullptr, n});
    }
    FCFD_CHECK_LAUNCH();
    return c.lv[0].y;
}

struct ZOut27 {
    const double* y; double* z; const DevState* st; int neumann;
    __device__ void operator()(int k, double (&)[1]) const { z[k] = y[k] - (neumann ? st->mg[0][MG_Y] : 0.0); }
};

void vcycle_h27_api(Ctx& c, const double* r, double* z) {
    FCFD_REQUIRE(c.precond == FCFD_PC_MULTIGRID, "context was not created with the multigrid preconditioner");
    const int nn = c.g3.nn;
    struct SumBody {
        const double* r;
        __device__ void operator()(int k, double (&acc)[1]) const { acc[0] += r[k]; }
    };
    launch_node<1>(c, nn, nullptr, SumBody{r}, MeanFin{&c.st->mg[0][MG_R], nullptr, nn});
    const double* y = vcycle_h27(c, r, nullptr);
    launch_node<0>(c, nn, nullptr, ZOut27{y, z, c.st, !c.set_zeroP}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

static const double* precondition_h27(Ctx& c, const int* active) {
    const int nn = c.g3.nn;
    const double* Le = c.eh27_dev->Le;
    switch (c.precond) {
        case FCFD_PC_NONE: return c.c_r;
        case FCFD_PC_JACOBI:
            launch_node<0>(c, nn, active, JacobiPrec{c.c_r, c.lv[0].Ld, c.lv[0].y}, NoFin{});
            return c.lv[0].y;
        case FCFD_PC_APPROX_INV:
            launch_ctile27<0>(c, c.g3, active, ApproxInv27a{c.g3, Le, c.c_r, c.lv[0].Ld, c.ppin, c.lv[0].x},
                             NoFin{});
            launch_ctile27<0>(c, c.g3, active,
                             ApproxInv27b{c.g3, Le, c.c_r, c.lv[0].Ld, c.ppin, c.lv[0].x, c.lv[0].y}, NoFin{});
            return c.lv[0].y;
        case FCFD_PC_MULTIGRID: return vcycle_h27(c, c.c_r, active);
    }
    throw Error(FCFD_ERR_UNSUPPORTED, "unknown hex27 pressure preconditioner");
}

// PCG (jax_iterative_solvers.py:87-120); solution in c.c_x
static int pcg_h27(Ctx& c, const double* b, const double* bmean, const double* x0in, int zero_pins, double tol,
                   int maxit, bool* converged) {
    const Grid3D& g = c.g3;
    const int nn = g.nn;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    const double* Le = c.eh27_dev->Le;
    launch_ctile27<3>(c, g, nullptr, CgInit27{g, Le, pin, b, bmean, x0in, zero_pins, c.c_x, c.c_r},
                     CgInitFin{c.st, tol, maxit, nn, 0});
    const int* act = &c.st->c_active;
    const int neu = c.precond == FCFD_PC_MULTIGRID && !c.set_zeroP;
    const double* y0 = precondition_h27(c, act);
    launch_node<1>(c, nn, act, CgGamma0<double>{c.c_r, y0, c.c_z[0], c.c_p[0], c.st, neu}, CgGamma0Fin{c.st});
    FCFD_CHECK_LAUNCH();
    const double* Ld = c.lv[0].Ld;
    auto pair = [&]() {
        for (int o = 0; o < 2; ++o) {
            const int nw = o ^ 1;
            launch_btile27<1>(c, g, act, CgK1_27{g, Le, pin, c.c_z[o], c.c_p[o], c.c_p[nw], c.c_Ap, c.st},
                             CgK1Fin{c.st});
            if (c.precond == FCFD_PC_JACOBI || c.precond == FCFD_PC_NONE) {
                launch_node<2>(c, nn, act,
                               CgK23{c.c_p[nw], c.c_Ap, c.precond == FCFD_PC_JACOBI ? Ld : nullptr, c.c_x, c.c_r,
                                     c.c_z[nw], c.st},
                               CgK23Fin{c.st});
            } else if (c.precond == FCFD_PC_APPROX_INV && c.coarse_inv32 == nullptr) {
                // the x / r update rides on stage 1 (its r.r partials in slot 2) and the z-step on
                // stage 2; r alternates between c_r (parity 0) and c_r2
                const double* rin = o == 0 ? c.c_r : c.c_r2;
                double* rout = o == 0 ? c.c_r2 : c.c_r;
                const int ntx = (g.nnx + C27X - 1) / C27X, nty = (g.nny + C27Y - 1) / C27Y,
                          ntz = (g.nnz + C27Z - 1) / C27Z;
                const int nbt = ntx * nty * ntz;
                if (nbt > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
                k_ctile27<1><<<nbt, C27T, 0, c.stream>>>(
                    g, ntx, nty, act,
                    ApproxInv27aK2{g, Le, rin, c.c_Ap, c.c_p[nw], c.c_x, rout, Ld, c.ppin, c.lv[0].x, c.st},
                    c.partials + 2 * (size_t)c.pstride, c.pstride);
                k_ctile27<2><<<nbt, C27T, 0, c.stream>>>(
                    g, ntx, nty, act, ApproxInv27bK3{g, Le, rout, Ld, c.ppin, c.lv[0].x, c.c_z[nw]}, c.partials,
                    c.pstride);
                finalize2_kernel<2, 1><<<1, FIN_BLOCK, 0, c.stream>>>(nbt, nbt, act, CgK3RFin{c.st, 0}, c.partials,
                                                                      c.pstride);
            } else if (neu) {
                launch_node<2>(c, nn, act, CgK2{c.c_p[nw], c.c_Ap, c.c_x, c.c_r, c.st}, CgK2Fin{c.st, nn});
                const double* y = precondition_h27(c, act);
                const int flex = c.coarse_inv32 != nullptr;
                launch_node<2>(c, nn, act, CgK3<double>{c.c_r, y, c.c_z[nw], c.c_z[o], c.st, neu, flex},
                               CgK3Fin{c.st, flex});
            } else {
                // CgK2's partials of r.r and sum r in slots 2, 3, taken by the z-step's finalize
                // (CgK3RFin: CgK2Fin's c_rr, then CgK3Fin); without the pure-Neumann V-cycle
                // nothing reads them before
                const int nb2 = blocks_for(nn);
                node_kernel<2><<<nb2, BLOCK, 0, c.stream>>>(nn, act, CgK2{c.c_p[nw], c.c_Ap, c.c_x, c.c_r, c.st},
                                                             c.partials + 2 * (size_t)c.pstride, c.pstride);
                const double* y = precondition_h27(c, act);
                const int flex = c.coarse_inv32 != nullptr;
                node_kernel<2><<<nb2, BLOCK, 0, c.stream>>>(
                    nn, act, CgK3<double>{c.c_r, y, c.c_z[nw], c.c_z[o], c.st, 0, flex}, c.partials, c.pstride);
                finalize2_kernel<2, 1><<<1, FIN_BLOCK, 0, c.stream>>>(nb2, nb2, act, CgK3RFin{c.st, flex},
                                                                      c.partials, c.pstride);
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

int pressure_cg_h27(Ctx& c, const double* b, double* x, double tol, int maxit, bool* converged) {
    const int it = pcg_h27(c, b, nullptr, x, 0, tol, maxit, converged);
    FCFD_HIP(hipMemcpyAsync(x, c.c_x, sizeof(double) * c.g3.nn, hipMemcpyDeviceToDevice, c.stream));
    return it;
}

int solve_pressure_h27(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                       double* P_out) {
    const int nn = c.g3.nn;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    launch_ctile27<1>(c, c.g3, nullptr, Prhs27Body{c.g3, c.eh27_dev, U_star, F_p, pin, c.c_rhs, rho / dt},
                     MeanFin{&c.st->mean_rhs, nullptr, nn});
    const int maxit = c.maxit_p > 0 ? c.maxit_p : 1000;
    bool conv;
    const int it = pcg_h27(c, c.c_rhs, c.set_zeroP ? nullptr : &c.st->mean_rhs, P_n, c.set_zeroP, c.tol, maxit, &conv);
    if (!c.set_zeroP) launch_node<1>(c, nn, nullptr, SumXBody{c.c_x}, MeanFin{&c.st->mean_p, nullptr, nn});
    launch_node<0>(c, nn, nullptr, PoutBody{c.c_x, pin, P_out, c.st, !c.set_zeroP}, NoFin{});
    FCFD_CHECK_LAUNCH();
    return it;
}

// ------------------------------------------------------------------------------ mass matvec
struct Mass27Body {
    Grid3D g; const ElemConstH27* ec; const double* x; double* y;
    __device__ void operator()(int n, double (&)[1]) const {
        Inc27 s;
        node_inc_h27(g, n, s);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double* xc = x + (size_t)d * g.nn;
            y[n + (size_t)d * g.nn] = h27_const(g, s, ec->Me, [&](int id) { return xc[id]; });
        }
    }
};
void mass_matvec_h27(Ctx& c, const double* x, double* y) {
    launch_node<0>(c, c.g3.nn, nullptr, Mass27Body{c.g3, c.eh27_dev, x, y}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ boundary integral
// F_p[n] = sum over the node's Dirichlet-face slots (bconn order) of sum_b W_f[a][b] (n_f . U_d(b)),
// W_f[a][b] = sum_q w_q det J_q N_qa N_qb of the quad9 face mapped from its (z, y) coordinates
__global__ void k_boundary_integral_h27(int nn, const int* __restrict__ off, const int* __restrict__ slot,
                                        const double* __restrict__ W, const double* __restrict__ nv,
                                        const int* __restrict__ ui, const double* __restrict__ U_d, double* F_p) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    double acc = 0.0;
    for (int k = off[n]; k < off[n + 1]; ++k) {
        const int f = slot[k] / 9, a = slot[k] - 9 * f;
        double t = 0.0;
        for (int b = 0; b < 9; ++b) {
            const int* u = ui + 27 * (size_t)f + 3 * b;
            t += W[81 * (size_t)f + a * 9 + b] *
                 (nv[3 * f] * U_d[u[0]] + nv[3 * f + 1] * U_d[u[1]] + nv[3 * f + 2] * U_d[u[2]]);
        }
        acc += t;
    }
    F_p[n] = acc;
}

void boundary_integral_h27(Ctx& c, const double* U_d, double* F_p) {
    const int nn = c.g3.nn;
    if (!c.fb_off) {
        launch_fill(c, F_p, 0.0, nn);
        return;
    }
    k_boundary_integral_h27<<<(nn + 255) / 256, 256, 0, c.stream>>>(nn, c.fb_off, c.fb_edge, c.fb_w, c.fb_geom,
                                                                     c.fb_ui, U_d, F_p);
    FCFD_CHECK_LAUNCH();
}

// quad9 3x3 Gauss weights of a face from the (z, y) coordinates of its nodes (quad9 local order):
// W[a][b] = sum_q w_q det J_q N_qa N_qb
static void face_weights9(const double (&P)[9][2], double (&W)[81]) {
    const double g = std::sqrt(3.0 / 5.0);
    const double x1[3] = {-g, 0.0, g}, w1[3] = {5.0 / 9, 8.0 / 9, 5.0 / 9};
    const int px[9] = {0, 1, 2, 2, 2, 1, 0, 0, 1}, py[9] = {0, 0, 0, 1, 2, 2, 2, 1, 1};
    auto L = [](double t, int a) { return a == 0 ? t * (t - 1) / 2 : (a == 1 ? 1 - t * t : t * (t + 1) / 2); };
    auto dL = [](double t, int a) { return a == 0 ? (2 * t - 1) / 2 : (a == 1 ? -2 * t : (2 * t + 1) / 2); };
    for (int k = 0; k < 81; ++k) W[k] = 0.0;
    for (int qe = 0; qe < 3; ++qe)
        for (int qx = 0; qx < 3; ++qx) {
            const double xi = x1[qx], eta = x1[qe];
            double N[9], dxi[9], deta[9];
            double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
            for (int b = 0; b < 9; ++b) {
                N[b] = L(xi, px[b]) * L(eta, py[b]);
                dxi[b] = dL(xi, px[b]) * L(eta, py[b]);
                deta[b] = L(xi, px[b]) * dL(eta, py[b]);
                J00 += dxi[b] * P[b][0]; J01 += dxi[b] * P[b][1];
                J10 += deta[b] * P[b][0]; J11 += deta[b] * P[b][1];
            }
            const double w = w1[qx] * w1[qe] * (J00 * J11 - J01 * J10);
            for (int a = 0; a < 9; ++a)
                for (int b = 0; b < 9; ++b) W[a * 9 + b] += w * N[a] * N[b];
        }
}

// ------------------------------------------------------------------------------ host set-up
template <class T>
static void upload27(const std::vector<T>& h, T* d) {
    FCFD_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

// out[n] = sum over the incident elements of per_local[the node's local index], element order
static void node_sums_h27(const Grid3D& g, const double (&per_local)[27], std::vector<double>& out) {
    out.assign(g.nn, 0.0);
    for (int ke = 0; ke < g.nez; ++ke)
        for (int je = 0; je < g.ney; ++je)
            for (int ie = 0; ie < g.nex; ++ie)
                for (int b = 0; b < 27; ++b)
                    out[(size_t)(2 * ke + h27z(b)) * g.nxy + (2 * je + h27y(b)) * g.nnx + 2 * ie + h27x(b)] +=
                        per_local[b];
}

void build_ctx_h27(Ctx& c, const fcfd_mesh_desc* m, const fcfd_solver_desc* s) {
    FCFD_REQUIRE(m->nex >= 1 && m->ney >= 1 && m->nez >= 1, "bad element counts");
    FCFD_REQUIRE(!m->periodic, "3D periodic meshes are not supported (the reference rejects them)");
    FCFD_REQUIRE(c.ce_format == FCFD_CE_F64 || c.ce_format == FCFD_CE_F32, "hex27 Ce formats: FP64 or FP32");
    FCFD_REQUIRE((int64_t)(2 * m->nex + 1) * (2 * m->ney + 1) * (2 * m->nez + 1) * 3 < (int64_t)1 << 31,
                 "grid too large");
    c.dim = 3;
    c.order = 2;
    c.g3 = make_grid27(m->nex, m->ney, m->nez);
    const Grid3D& g = c.g3;
    c.g = Grid2D{0, 0, 0, 0, g.nn, g.ne, 0};
    c.lx = m->lx; c.ly = m->ly; c.lz = m->lz;
    const int nn = g.nn;
    c.eh27.resize(1);
    ElemConstH27& ec = c.eh27[0];
    hex27_constants(c.lx / g.nex, c.ly / g.ney, c.lz / g.nez, ec);
    c.eh27_dev = dalloc<ElemConstH27>(c, 1);
    FCFD_HIP(hipMemcpy(c.eh27_dev, &ec, sizeof(ElemConstH27), hipMemcpyHostToDevice));
    {
        double rs[27];
        for (int a = 0; a < 27; ++a) {
            rs[a] = 0.0;
            for (int b = 0; b < 27; ++b) rs[a] += ec.Me[a * 27 + b];
        }
        std::vector<double> M;
        node_sums_h27(g, rs, M);
        M.resize(3 * (size_t)nn);
        std::copy(M.begin(), M.begin() + nn, M.begin() + nn);
        std::copy(M.begin(), M.begin() + nn, M.begin() + 2 * (size_t)nn);
        c.Mlump = dalloc<double>(c, 3 * (size_t)nn);
        upload27(M, c.Mlump);
    }
    // levels: hex27 meshes of nex >> l elements (level 0 always exists: jacobi / approx_inv use it)
    const int nlev = c.precond == FCFD_PC_MULTIGRID ? s->mg_levels : 1;
    FCFD_REQUIRE(nlev >= 1 && nlev <= MAX_LEVELS, "bad multigrid level count");
    FCFD_REQUIRE((g.nex >> (nlev - 1)) << (nlev - 1) == g.nex && (g.ney >> (nlev - 1)) << (nlev - 1) == g.ney &&
                     (g.nez >> (nlev - 1)) << (nlev - 1) == g.nez && (g.nex >> (nlev - 1)) >= 1 &&
                     (g.ney >> (nlev - 1)) >= 1 && (g.nez >> (nlev - 1)) >= 1,
                 "3D multigrid: every element count must halve exactly at each level");
    c.nlev = nlev;
    c.lv.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
        Level& L = c.lv[l];
        L.g3 = make_grid27(g.nex >> l, g.ney >> l, g.nez >> l);
        L.g = Grid2D{0, 0, 0, 0, L.g3.nn, L.g3.ne, 0};
        std::vector<ElemConstH27> el(1);
        if (l == 0) el[0] = ec;
        else hex27_constants(c.lx / L.g3.nex, c.ly / L.g3.ney, c.lz / L.g3.nez, el[0]);
        L.eh27_dev = dalloc<ElemConstH27>(c, 1);
        FCFD_HIP(hipMemcpy(L.eh27_dev, el.data(), sizeof(ElemConstH27), hipMemcpyHostToDevice));
        double dg[27];
        for (int a = 0; a < 27; ++a) dg[a] = el[0].Le[a * 27 + a];
        std::vector<double> Ld;
        node_sums_h27(L.g3, dg, Ld);
        const int n = L.g3.nn;
        L.Ld = dalloc<double>(c, n);
        upload27(Ld, L.Ld);
        if (l > 0) L.r = dalloc<double>(c, n);
        L.x = dalloc<double>(c, n);
        L.y = dalloc<double>(c, n);
        if (nlev > 1 || c.precond == FCFD_PC_MULTIGRID) L.f = dalloc<double>(c, n);
        if (c.set_zeroP && c.precond == FCFD_PC_MULTIGRID) {
            std::vector<uint8_t> pin(n, 0);
            if (s->mg_gamma_p && s->mg_gamma_p_off) {
                for (int64_t k = s->mg_gamma_p_off[l]; k < s->mg_gamma_p_off[l + 1]; ++k) {
                    const int64_t v = s->mg_gamma_p[k];
                    FCFD_REQUIRE(v >= 0 && v < n, "multigrid pin out of range");
                    pin[v] = 1;
                }
            } else {
                pin[0] = 1;
            }
            L.pin = dalloc<uint8_t>(c, n);
            upload27(pin, L.pin);
        }
    }
    if (c.precond == FCFD_PC_MULTIGRID) {
        const int ncoarse = c.lv[nlev - 1].g3.nn;
        FCFD_REQUIRE(!c.set_zeroP || (size_t)ncoarse * sizeof(double) <= 64 * 1024,
                     "3D multigrid: the pinned coarse grid must have at most 8192 nodes");
        c.cc_r = dalloc<double>(c, ncoarse);
        c.cc_Ap = dalloc<double>(c, ncoarse);
        // mixed form, pinned pressure: exact inverse of the pinned coarse hex27 operator (FP32)
        if (c.precision == FCFD_PREC_MIXED && c.set_zeroP && nlev > 1 && ncoarse <= COARSE3_DENSE_MAX) {
            const Grid3D& gc = c.lv[nlev - 1].g3;
            ElemConstH27 ecc;
            hex27_constants(c.lx / gc.nex, c.ly / gc.ney, c.lz / gc.nez, ecc);
            std::vector<double> A((size_t)ncoarse * ncoarse, 0.0), inv;
            for (int ke = 0; ke < gc.nez; ++ke)
                for (int je = 0; je < gc.ney; ++je)
                    for (int ie = 0; ie < gc.nex; ++ie) {
                        const int n0 = 2 * ke * gc.nxy + 2 * je * gc.nnx + 2 * ie;
                        for (int a = 0; a < 27; ++a) {
                            const size_t ra = (size_t)(n0 + h27z(a) * gc.nxy + h27y(a) * gc.nnx + h27x(a)) * ncoarse;
                            for (int b = 0; b < 27; ++b)
                                A[ra + n0 + h27z(b) * gc.nxy + h27y(b) * gc.nnx + h27x(b)] += ecc.Le[a * 27 + b];
                        }
                    }
            std::vector<uint8_t> pin(ncoarse);
            FCFD_HIP(hipMemcpy(pin.data(), c.lv[nlev - 1].pin, ncoarse, hipMemcpyDeviceToHost));
            std::vector<int64_t> pins;
            for (int k = 0; k < ncoarse; ++k)
                if (pin[k]) pins.push_back(k);
            pin_identity(A, ncoarse, pins);
            if (spd_inverse_banded(A, ncoarse, 2 * gc.nxy + 2 * gc.nnx + 2, inv, true)) {
                std::vector<float> inv32(inv.begin(), inv.end());
                c.coarse_inv32 = dalloc<float>(c, inv32.size());
                upload27(inv32, c.coarse_inv32);
            }
        }
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
        upload27(pin, c.ppin);
    }
    std::vector<int64_t> gd;
    if (s->gamma_d && s->n_gamma_d > 0) {
        gd.assign(s->gamma_d, s->gamma_d + s->n_gamma_d);
        for (size_t k = 0; k < gd.size(); ++k) {
            FCFD_REQUIRE(gd[k] >= 0 && gd[k] < 3 * (int64_t)nn, "gamma_d out of range");
            FCFD_REQUIRE(k == 0 || gd[k] > gd[k - 1], "gamma_d must be sorted and unique");
        }
        c.n_gamma_d = (int)gd.size();
        std::vector<uint8_t> mask(3 * (size_t)nn, 0);
        for (auto v : gd) mask[v] = 1;
        c.dmask = dalloc<uint8_t>(c, 3 * (size_t)nn);
        upload27(mask, c.dmask);
        c.gamma_d_dev = dalloc<int64_t>(c, gd.size());
        upload27(gd, c.gamma_d_dev);
        c.ud_full = dalloc<double>(c, 3 * (size_t)nn);
        FCFD_HIP(hipMemset(c.ud_full, 0, sizeof(double) * 3 * nn));
    }
    // Dirichlet faces (9 nodes each) -> per-node gather lists in bconn order
    if (s->bconn && s->nvec && s->n_bedge > 0 && !gd.empty()) {
        const int nf = (int)s->n_bedge;
        c.n_bedge = nf;
        std::vector<std::vector<int>> lists(nn);
        std::vector<double> W(81 * (size_t)nf), geom(3 * (size_t)nf);
        std::vector<int> ui(27 * (size_t)nf);
        const double hy = c.ly / g.ney / 2, hz = c.lz / g.nez / 2;
        auto sorted_index = [&](int64_t v) { return (int)(std::lower_bound(gd.begin(), gd.end(), v) - gd.begin()); };
        for (int f = 0; f < nf; ++f) {
            double P[9][2];
            for (int a = 0; a < 9; ++a) {
                const int64_t v = s->bconn[9 * (size_t)f + a];
                FCFD_REQUIRE(v >= 0 && v < nn, "bconn out of range");
                const int kk = (int)(v / g.nxy), r = (int)(v % g.nxy);
                P[a][0] = kk * hz;            // z
                P[a][1] = (r / g.nnx) * hy;   // y
                for (int d = 0; d < 3; ++d) {
                    const int idx = sorted_index(v + (int64_t)d * nn);
                    FCFD_REQUIRE(idx < (int)gd.size() && gd[idx] == v + (int64_t)d * nn,
                                 "boundary node is not a Dirichlet DOF");
                    ui[27 * (size_t)f + 3 * a + d] = idx;
                }
                lists[v].push_back(f * 9 + a);
            }
            double w[81];
            face_weights9(P, w);
            for (int k = 0; k < 81; ++k) W[81 * (size_t)f + k] = w[k];
            for (int d = 0; d < 3; ++d) geom[3 * (size_t)f + d] = (double)s->nvec[3 * (size_t)f + d];
        }
        std::vector<int> off(nn + 1, 0), slots;
        for (int n = 0; n < nn; ++n) {
            off[n + 1] = off[n] + (int)lists[n].size();
            slots.insert(slots.end(), lists[n].begin(), lists[n].end());
        }
        c.fb_off = dalloc<int>(c, nn + 1);
        upload27(off, c.fb_off);
        c.fb_edge = dalloc<int>(c, slots.size());
        upload27(slots, c.fb_edge);
        c.fb_w = dalloc<double>(c, W.size());
        upload27(W, c.fb_w);
        c.fb_geom = dalloc<double>(c, geom.size());
        upload27(geom, c.fb_geom);
        c.fb_ui = dalloc<int>(c, ui.size());
        upload27(ui, c.fb_ui);
    }
    const size_t ne = g.ne;
    if (c.ce_format == FCFD_CE_F64) c.ce = dalloc<double>(c, 27 * H27_ROW * ne);
    else c.ce = dalloc<float>(c, 27 * H27_ROW * ne);
    const size_t n3 = 3 * (size_t)nn;
    c.m_r = dalloc<double>(c, n3);
    c.m_rhat = dalloc<double>(c, n3);
    for (int k = 0; k < 2; ++k) {
        c.m_p[k] = dalloc<double>(c, n3);
        c.m_q[k] = dalloc<double>(c, n3);
        c.c_p[k] = dalloc<double>(c, nn);
        c.c_z[k] = dalloc<double>(c, nn);
    }
    c.m_s = dalloc<double>(c, n3);
    c.m_t = dalloc<double>(c, n3);
    c.m_x = dalloc<double>(c, n3);
    c.c_r = dalloc<double>(c, nn);
    if (c.precond == FCFD_PC_APPROX_INV) c.c_r2 = dalloc<double>(c, nn);  // the fused x / r update's r_out
    c.c_Ap = dalloc<double>(c, nn);
    c.c_x = dalloc<double>(c, nn);
    c.c_rhs = dalloc<double>(c, nn);
    c.ustar = dalloc<double>(c, n3);
    c.pstride = blocks_for((int)n3) + 1;
    c.partials = dalloc<double>(c, (size_t)MAX_RED * c.pstride);
    c.st = dalloc<DevState>(c, 1);
    FCFD_HIP(hipMemset(c.st, 0, sizeof(DevState)));
    FCFD_HIP(hipHostMalloc((void**)&c.h_pinned, 64, hipHostMallocDefault));
    FCFD_HIP(hipDeviceSynchronize());
}

// ------------------------------------------------------------------------------ the step
void four_step_h27(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* U_d,
                   const double* F_p, const double* F, double* U, double* P, int* it) {
    scatter_ud(c, U_d);
    it[0] = solve_momentum_h27(c, U_n, P_n, rho, mu, dt, F, c.m_x);
    launch_ctile27<0>(c, c.g3, nullptr,
                     Proj27Body{c.g3, c.eh27_dev, c.m_x, P_n, c.Mlump, nullptr, nullptr, c.ustar, dt / rho}, NoFin{});
    FCFD_CHECK_LAUNCH();
    it[1] = solve_pressure_h27(c, c.ustar, P_n, F_p, rho, dt, P);
    launch_ctile27<0>(c, c.g3, nullptr,
                     Proj27Body{c.g3, c.eh27_dev, c.ustar, P, c.Mlump, c.ud_full, c.dmask, U, -dt / rho}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

}  // namespace fcfd
