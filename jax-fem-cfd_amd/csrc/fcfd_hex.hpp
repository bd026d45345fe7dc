// Structured Q1 hexahedral (hex8) grids: index math, element constants, 27-point node
// neighbourhoods and row-blocked 8 x 8 element-matrix storage.
//
// Mesh of create_mesh_3d (jax_fem_cfd/setup/regular_mesh.py:45-105): node (i, j, k) ->
// k*nnx*nny + j*nnx + i at (i dx, j dy, k dz); element (ie, je, ke) -> (ke*ney + je)*nex + ie
// with corners [quad at ke] ++ [quad at ke+1], quad = (ie,je), (ie+1,je), (ie+1,je+1), (ie,je+1).
// The reference has no periodic 3D mesh (simulation.py:77-78 rejects it), so neither has this.
#pragma once
#include "fcfd_common.hpp"

namespace fcfd {

struct Grid3D {
    int nex, ney, nez, nnx, nny, nnz, nxy, nn, ne;
};

inline Grid3D make_grid3(int nex, int ney, int nez) {
    Grid3D g;
    g.nex = nex; g.ney = ney; g.nez = nez;
    g.nnx = nex + 1; g.nny = ney + 1; g.nnz = nez + 1;
    g.nxy = g.nnx * g.nny;
    g.nn = g.nxy * g.nnz;
    g.ne = nex * ney * nez;
    return g;
}

// hex8 2x2x2 Gauss data and element constants of a box element (shape_functions.py:47-86,
// element_calc.py:9-40): [q][I] for N and the physical derivatives, Me = sum_q w N N^T,
// Le = sum_q w grad N^T grad N, G[d] = sum_q w N (d_d N)^T.
struct ElemConst3 {
    double N[64], dN[3][64], wq[8];
    double Me[64], Le[64], G[3][64];
    double Leoff[64];  // Le with its diagonal zeroed (the Jacobi smoother's off-diagonal part)
};

// Triquadratic hex27 constants of a box element (3x3x3 Gauss, q = (qz*3 + qy)*3 + qx; local node
// b = (c*3 + b_y)*3 + a at (a, b_y, c) in {0,1,2}^3), [q][b] / [a][b] layouts as ElemConst3.
struct ElemConstH27 {
    double N[729], dN[3][729], wq[27];
    double Me[729], Le[729], Leoff[729], G[3][729];
    double WN[729];  // [a][q] = w_q N_qa: the test-function weights of the Ce build
};

// Local corner numbering: corner b sits at (hbx, hby, hbz) in {0,1}^3 of its element.
__host__ __device__ constexpr int hbx(int b) { return ((b & 3) == 1 || (b & 3) == 2) ? 1 : 0; }
__host__ __device__ constexpr int hby(int b) { return (b & 3) >= 2 ? 1 : 0; }
__host__ __device__ constexpr int hbz(int b) { return b >> 2; }
__host__ __device__ constexpr int hloc(int x, int y, int z) { return (y == 0 ? x : 3 - x) + 4 * z; }

// Incident element k = kx + 2 ky + 4 kz of node (i, j, k) is element (i-1+kx, j-1+ky, k-1+kz)
// (ascending element id in k, the order segment_sum accumulates in).  The node is its corner
// k3Row(k); corner b of that element is the node's 27-point stencil slot k3Slot(k, b)
// ((dz+1)*3 + (dy+1))*3 + (dx+1).
__host__ __device__ constexpr int k3Row(int k) { return hloc(1 - (k & 1), 1 - ((k >> 1) & 1), 1 - (k >> 2)); }
__host__ __device__ constexpr int k3Slot(int k, int b) {
    return ((hbz(b) + (k >> 2)) * 3 + (hby(b) + ((k >> 1) & 1))) * 3 + (hbx(b) + (k & 1));
}
constexpr int S3C = 13;  // the centre slot

// 27-point neighbourhood of a node: element ids clamped to 0 and node ids clamped to the node
// itself where missing, flagged, so every load issues unconditionally.
struct Nbr3 {
    int e[8];
    bool has[8];
    int nb[27];
};

__device__ __forceinline__ void node_nbr3(const Grid3D& g, int n, Nbr3& s) {
    const int k = n / g.nxy, r = n - k * g.nxy;
    const int j = r / g.nnx, i = r - j * g.nnx;
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int ii = i + dx - 1, jj = j + dy - 1, kk = k + dz - 1;
                const bool ok = ii >= 0 && ii < g.nnx && jj >= 0 && jj < g.nny && kk >= 0 && kk < g.nnz;
                s.nb[(dz * 3 + dy) * 3 + dx] = ok ? kk * g.nxy + jj * g.nnx + ii : n;
            }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int ie = i - 1 + (q & 1), je = j - 1 + ((q >> 1) & 1), ke = k - 1 + (q >> 2);
        const bool h = ie >= 0 && ie < g.nex && je >= 0 && je < g.ney && ke >= 0 && ke < g.nez;
        s.has[q] = h;
        s.e[q] = h ? (ke * g.ney + je) * g.nex + ie : 0;
    }
}

__device__ __forceinline__ void elem_nodes3(const Grid3D& g, int e, int (&c)[8]) {
    const int ke = e / (g.nex * g.ney), r = e - ke * g.nex * g.ney;
    const int je = r / g.nex, ie = r - je * g.nex;
    const int n0 = ke * g.nxy + je * g.nnx + ie;
#pragma unroll
    for (int b = 0; b < 8; ++b) c[b] = n0 + hbz(b) * g.nxy + hby(b) * g.nnx + hbx(b);
}

// 27-point gather (values of missing neighbours belong only to missing elements, which the
// element loops mask, so they are loaded from the clamped id and never used)
template <class F>
__device__ __forceinline__ void load27(const Nbr3& s, F f, double (&X)[27]) {
#pragma unroll
    for (int t = 0; t < 27; ++t) X[t] = f(s.nb[t]);
}

// y(node) = sum_k sum_b A[row_k][b] X[slot_k(b)] with one shared 8 x 8 matrix (TR: A^T)
template <bool TR = false>
__device__ __forceinline__ double ebe3(const Nbr3& s, const double* __restrict__ A, const double (&X)[27]) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int a = k3Row(k);
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b) t += (TR ? A[b * 8 + a] : A[a * 8 + b]) * X[k3Slot(k, b)];
        acc += s.has[k] ? t : 0.0;
    }
    return acc;
}


// incident elements of node n (ids clamped to 0 where missing)
__device__ __forceinline__ void node_elems3(const Grid3D& g, int n, Nbr3& s) {
    const int k = n / g.nxy, r = n - k * g.nxy;
    const int j = r / g.nnx, i = r - j * g.nnx;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int ie = i - 1 + (q & 1), je = j - 1 + ((q >> 1) & 1), ke = k - 1 + (q >> 2);
        const bool h = ie >= 0 && ie < g.nex && je >= 0 && je < g.ney && ke >= 0 && ke < g.nez;
        s.has[q] = h;
        s.e[q] = h ? (ke * g.ney + je) * g.nex + ie : 0;
    }
}

template <class SV>
__device__ __forceinline__ void view27(const SV& sv, int v, int n, double (&X)[27]) {
#pragma unroll
    for (int t = 0; t < 27; ++t) X[t] = sv.at(v, n, t);
}

// ---------------------------------------------------------------------------------------------
// LDS-staged 27-point kernels over box tiles.  A block owns S3T_X x S3T_Y x S3T_Z nodes (one per
// thread) and stages the box one node wider (18 x 6 x 6 = 2.5 staged values per node; the body
// forms the staged value, e.g. z + beta p, once per window node): every node then reads its
// stencil from LDS.  The 3 x 3 (dy, dz) neighbour rows of 256 consecutive nodes would stage 9
// values per node and measured 6-15% slower per 3D step; flatter boxes (64 x 2 x 2, 32 x 8 x 1,
// 32 x 4 x 2) were slower than 16 x 4 x 4 too.  Window nodes outside the grid stage 0: the
// bodies mask the elements a node lacks with selects, so the values at such stencil points never
// enter a result.  Each thread's view holds the LDS address of its stencil's first point, so the
// 27 reads are compile-time offsets.
// Body: static constexpr int NV; void stage(int id, double (&v)[NV]) const;
//       template <class SV> void operator()(int n, const SV& view, double (&acc)[NR]) const,
//       reading stencil slot t = ((dz+1)*3 + (dy+1))*3 + (dx+1) of value v as view.at(v, n, t).
constexpr int S3T_X = 16, S3T_Y = 4, S3T_Z = 4;
template <int NV, int WX, int WY, int WN>
struct StageView3T {
    const double* b;  // this thread's stencil point (dx, dy, dz) = (-1, -1, -1) in the window of value 0
    __device__ __forceinline__ double at(int v, int, int t) const {
        return b[v * WN + (t / 9) * (WX * WY) + ((t / 3) % 3) * WX + (t % 3)];
    }
};
template <int NR, class Body>
__global__ __launch_bounds__(BLOCK) void stencil3t_kernel(Grid3D g, int ntx, int nty, const int* __restrict__ active,
                                                         Body body, double* __restrict__ partials, int pstride) {
    constexpr int TX = S3T_X, TY = S3T_Y, TZ = S3T_Z;
    static_assert(TX * TY * TZ == BLOCK, "one node per thread");
    constexpr int NV = Body::NV, WX = TX + 2, WY = TY + 2, WZ = TZ + 2, WN = WX * WY * WZ;
    constexpr int NIT = (WN + BLOCK - 1) / BLOCK;
    __shared__ double sm[NV * WN];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int tz = lb / (ntx * nty), trem = lb - tz * (ntx * nty), ty = trem / ntx, tx = trem - ty * ntx;
    const int i0 = tx * TX - 1, j0 = ty * TY - 1, k0 = tz * TZ - 1;  // window origin
    // every staging load of the thread is issued before the first LDS store
    double v[NIT][NV];
    int at[NIT];
    bool inw[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int w = threadIdx.x + it * BLOCK;
        const int wz = w / (WX * WY), wr = w - wz * (WX * WY), wy = wr / WX, wx = wr - wy * WX;
        const int i = i0 + wx, j = j0 + wy, k = k0 + wz;
        const bool in = w < WN && i >= 0 && i < g.nnx && j >= 0 && j < g.nny && k >= 0 && k < g.nnz;
        at[it] = w < WN ? w : -1;
        inw[it] = in;
        body.stage(in ? (k * g.nny + j) * g.nnx + i : 0, v[it]);
    }
    // the gate's load goes out behind the staging loads (its kernel-argument load then overlaps them)
    const int go = gate_read(active);
#pragma unroll
    for (int it = 0; it < NIT; ++it)
        if (at[it] >= 0) {
#pragma unroll
            for (int q = 0; q < NV; ++q) sm[q * WN + at[it]] = inw[it] ? v[it][q] : 0.0;
        }
    __syncthreads();
    if (!gate_open(go)) return;
    double acc[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < (NR > 0 ? NR : 1); ++r) acc[r] = 0.0;
    const int lx = threadIdx.x % TX, ly = (threadIdx.x / TX) % TY, lz = threadIdx.x / (TX * TY);
    const int i = i0 + 1 + lx, j = j0 + 1 + ly, k = k0 + 1 + lz;
    if (i < g.nnx && j < g.nny && k < g.nnz)
        body((k * g.nny + j) * g.nnx + i, StageView3T<NV, WX, WY, WN>{sm + (lz * WY + ly) * WX + lx}, acc);
    if constexpr (NR > 0) reduce_epilogue<NR>(acc, partials, pstride, lb);
}
// the grid of box tiles of a level
inline dim3 s3t_tiles(const Grid3D& g) {
    return dim3((g.nnx + S3T_X - 1) / S3T_X, (g.nny + S3T_Y - 1) / S3T_Y, (g.nnz + S3T_Z - 1) / S3T_Z);
}

// Row-blocked 8 x 8 storage: entry (a, b) of element e at ((a * ne + e) * 8 + b); a node reads
// its row of an incident element with contiguous vector loads (32 B FP32, 64 B FP64).
__host__ __device__ __forceinline__ size_t ce3_idx(int k, int e, int ne) {
    return ((size_t)(k >> 3) * ne + e) * 8 + (k & 7);
}
struct Ce3F64 {
    double* c; int ne;
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[8]) const {
        const double2* p = reinterpret_cast<const double2*>(c + ce3_idx(a * 8, e, ne));
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const double2 x = p[h];
            v[2 * h] = x.x; v[2 * h + 1] = x.y;
        }
    }
    __device__ __forceinline__ void st(int k, int e, double v) const { c[ce3_idx(k, e, ne)] = v; }
    static constexpr size_t bytes_per_elem = 64 * 8;
};
struct Ce3F32 {
    float* c; int ne;
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[8]) const {
        const float4* p = reinterpret_cast<const float4*>(c + ce3_idx(a * 8, e, ne));
        const float4 x = p[0], y = p[1];
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
    }
    __device__ __forceinline__ void st(int k, int e, double v) const { c[ce3_idx(k, e, ne)] = (float)v; }
    static constexpr size_t bytes_per_elem = 64 * 4;
};

}  // namespace fcfd
