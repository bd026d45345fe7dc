// Shared device infrastructure: structured-grid index math, deterministic
// two-stage reductions, XCD-aware block mapping, element-matrix storage.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstdio>
#include <string>

namespace fcfd {

constexpr int BLOCK = 256;         // threads per node/element kernel block (4 waves)
constexpr int MAX_BLOCKS = 2048;   // cap on grid size of reducing kernels
constexpr int NXCD = 8;            // MI355X: 8 XCDs, blocks dispatched round-robin

// ---------------------------------------------------------------------------------------------
// Structured Q1 grid.  Node (i, j) -> j*nnx + i; element (ie, je) -> je*nex + ie with corners
// [(ie,je), (ie+1,je), (ie+1,je+1), (ie,je+1)] (counter-clockwise).  Periodic grids have
// nnx = nex, nny = ney and wrap.
// ---------------------------------------------------------------------------------------------
struct Grid2D {
    int nex, ney, nnx, nny, nn, ne, periodic;
};

inline Grid2D make_grid(int nex, int ney, int periodic) {
    Grid2D g;
    g.nex = nex; g.ney = ney; g.periodic = periodic;
    g.nnx = periodic ? nex : nex + 1;
    g.nny = periodic ? ney : ney + 1;
    g.nn = g.nnx * g.nny;
    g.ne = nex * ney;
    return g;
}

// Logical block index from the hardware block index such that each XCD
// receives one contiguous range of logical blocks (neighbouring grid rows then
// share an L2).  Bijective on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    int x = b % NXCD, local = b / NXCD;
    int per = nb / NXCD, rem = nb % NXCD;
    return x * per + (x < rem ? x : rem) + local;
}

// Stencil context of one node: the four incident elements (element to the
// lower-left, lower-right, upper-left, upper-right of the node) with the node's
// local index in each (2, 3, 1, 0), and the 3x3 neighbour node ids.
struct NodeNbr {
    int i, j;
    int e[4];       // element ids (valid only where has[k])
    bool has[4];
    int nb[3][3];   // nb[dy+1][dx+1] node id, -1 when outside a non-periodic grid
};

__device__ __forceinline__ void node_nbr(const Grid2D& g, int n, NodeNbr& s) {
    int j = n / g.nnx, i = n - j * g.nnx;
    s.i = i; s.j = j;
    if (g.periodic) {
        int im = i == 0 ? g.nnx - 1 : i - 1, ip = i == g.nnx - 1 ? 0 : i + 1;
        int jm = j == 0 ? g.nny - 1 : j - 1, jp = j == g.nny - 1 ? 0 : j + 1;
        int xs[3] = {im, i, ip}, ys[3] = {jm, j, jp};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) s.nb[a][b] = ys[a] * g.nnx + xs[b];
        s.e[0] = jm * g.nex + im; s.e[1] = jm * g.nex + i;
        s.e[2] = j * g.nex + im;  s.e[3] = j * g.nex + i;
        s.has[0] = s.has[1] = s.has[2] = s.has[3] = true;
    } else {
        for (int a = 0; a < 3; ++a) {
            int jj = j + a - 1;
            for (int b = 0; b < 3; ++b) {
                int ii = i + b - 1;
                bool ok = (ii >= 0) && (ii < g.nnx) && (jj >= 0) && (jj < g.nny);
                s.nb[a][b] = ok ? jj * g.nnx + ii : -1;
            }
        }
        bool l = i >= 1, r = i < g.nex, d = j >= 1, u = j < g.ney;
        s.has[0] = l && d; s.has[1] = r && d; s.has[2] = l && u; s.has[3] = r && u;
        s.e[0] = (j - 1) * g.nex + (i - 1); s.e[1] = (j - 1) * g.nex + i;
        s.e[2] = j * g.nex + (i - 1);       s.e[3] = j * g.nex + i;
    }
}

// Local row of the node in incident element k (0..3 = element to the lower-left,
// lower-right, upper-left, upper-right of the node), and the 3x3 stencil slot
// [kCy][kCx] = [dy+1][dx+1] of that element's corner b.  constexpr so that
// fully unrolled loops fold them to immediates.
__host__ __device__ constexpr int kRow(int k) { return k == 0 ? 2 : (k == 1 ? 3 : (k == 2 ? 1 : 0)); }
__host__ __device__ constexpr int kCy(int k, int b) { return (k < 2 ? 0 : 1) + (b < 2 ? 0 : 1); }
__host__ __device__ constexpr int kCx(int k, int b) { return (k & 1) + ((b == 1 || b == 2) ? 1 : 0); }

// ---------------------------------------------------------------------------------------------
// Element constants (FP64, shared by every element of a uniform grid).
// ---------------------------------------------------------------------------------------------
struct ElemConst {
    double Me[16], Le[16], Leoff[16], Gx[16], Gy[16];
    double N[16], dNx[16], dNy[16], wq[4];  // quad4 2x2 Gauss data: [q][I]
};

// ---------------------------------------------------------------------------------------------
// Per-element matrix storage, entry-major: entry k = a*4 + b of element e at k*ne + e, so a
// wavefront reading one entry of 64 consecutive elements is one coalesced 512 B (FP64) read.
// ---------------------------------------------------------------------------------------------
struct CeF64 {
    double* c; int ne;
    __device__ __forceinline__ double ld(int k, int e) const { return c[(size_t)k * ne + e]; }
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        for (int k = 0; k < 16; ++k) c[(size_t)k * ne + e] = v[k];
    }
    static constexpr size_t bytes_per_elem = 16 * 8;
};
struct CeF32 {
    float* c; int ne;
    __device__ __forceinline__ double ld(int k, int e) const { return (double)c[(size_t)k * ne + e]; }
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        for (int k = 0; k < 16; ++k) c[(size_t)k * ne + e] = (float)v[k];
    }
    static constexpr size_t bytes_per_elem = 16 * 4;
};
// FP16 entries normalised by a per-element FP32 scale (max |entry|), 36 B / element
struct CeF16S {
    __half* c; float* s; int ne;
    __device__ __forceinline__ double ld(int k, int e) const {
        return (double)(__half2float(c[(size_t)k * ne + e]) * s[e]);
    }
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        double m = 0.0;
        for (int k = 0; k < 16; ++k) m = fmax(m, fabs(v[k]));
        float sc = m > 0.0 ? (float)m : 1.0f;
        s[e] = sc;
        for (int k = 0; k < 16; ++k) c[(size_t)k * ne + e] = __float2half((float)(v[k] / (double)sc));
    }
    static constexpr size_t bytes_per_elem = 16 * 2 + 4;
};
struct CeBF16S {
    __hip_bfloat16* c; float* s; int ne;
    __device__ __forceinline__ double ld(int k, int e) const {
        return (double)(__bfloat162float(c[(size_t)k * ne + e]) * s[e]);
    }
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        double m = 0.0;
        for (int k = 0; k < 16; ++k) m = fmax(m, fabs(v[k]));
        float sc = m > 0.0 ? (float)m : 1.0f;
        s[e] = sc;
        for (int k = 0; k < 16; ++k) c[(size_t)k * ne + e] = __float2bfloat16((float)(v[k] / (double)sc));
    }
    static constexpr size_t bytes_per_elem = 16 * 2 + 4;
};

// ---------------------------------------------------------------------------------------------
// Deterministic reductions.  Each reducing kernel accumulates NR per-thread sums, reduces them
// per block in a fixed tree, writes one partial per block, and the last block to finish sums
// the partials in fixed order and runs the kernel's finalize step (device scalar updates).
// No float atomics anywhere: results are bitwise reproducible run to run.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// All threads return the block totals (any block size that is a multiple of 64, up to 1024).
template <int NR>
__device__ __forceinline__ void block_sum(double (&v)[NR]) {
    __shared__ double red[NR][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double s = wave_sum(v[r]);
        if (lane == 0) red[r][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double s = 0.0;
        for (int k = 0; k < nw; ++k) s += red[r][k];
        v[r] = s;
    }
    __syncthreads();
}

// Publish this block's partials; returns true (uniformly) in the last block, which then sees
// every block's partials.
template <int NR>
__device__ __forceinline__ bool publish_partials(const double (&v)[NR], double* partials,
                                                 unsigned* counter, int lb, int nb) {
    __shared__ int last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
            __hip_atomic_store(&partials[r * MAX_BLOCKS + lb], v[r], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        unsigned prev = atomicAdd(counter, 1u);
        last = (prev == (unsigned)(nb - 1));
        if (last) atomicExch(counter, 0u);
    }
    __syncthreads();
    bool l = last;
    if (l) __threadfence();
    return l;
}

// Called by every thread of the last block: fixed-order totals of all partials.
template <int NR>
__device__ __forceinline__ void gather_partials(double (&tot)[NR], const double* partials, int nb) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double s = 0.0;
        for (int k = threadIdx.x; k < nb; k += BLOCK)
            s += __hip_atomic_load(&partials[r * MAX_BLOCKS + k], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        tot[r] = s;
    }
    block_sum<NR>(tot);
}

// Work split of a node/element range over a capped grid: logical block lb owns the contiguous
// range [lo, hi).
struct Range {
    int lo, hi;
};
__device__ __forceinline__ Range block_range(int n, int lb, int nb) {
    int chunk = (n + nb - 1) / nb;
    chunk = (chunk + BLOCK - 1) / BLOCK * BLOCK;
    int lo = lb * chunk;
    int hi = lo + chunk;
    if (hi > n) hi = n;
    if (lo > n) lo = n;
    return {lo, hi};
}

inline int grid_for(int n) {
    int b = (n + BLOCK - 1) / BLOCK;
    if (b > MAX_BLOCKS) b = MAX_BLOCKS;
    if (b < 1) b = 1;
    return b;
}

// Generic node kernel with NR fused reductions.  Body: void operator()(int n, double (&acc)[NR])
// Fin: void operator()(const double (&tot)[NR]) run by thread 0 of the last block.
// `active` (device flag, may be null) lets a whole chunk of solver iterations be enqueued
// ahead: once the solver has converged every later kernel returns at entry.
template <int NR, class Body, class Fin>
__global__ __launch_bounds__(BLOCK) void node_kernel(int n, const int* __restrict__ active, Body body,
                                                    Fin fin, double* partials, unsigned* counter) {
    if (active && *active == 0) return;
    const int nb = gridDim.x;
    const int lb = xcd_block(blockIdx.x, nb);
    Range rg = block_range(n, lb, nb);
    double acc[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < (NR > 0 ? NR : 1); ++r) acc[r] = 0.0;
    for (int k = rg.lo + threadIdx.x; k < rg.hi; k += BLOCK) body(k, acc);
    if constexpr (NR > 0) {
        block_sum<NR>(acc);
        if (publish_partials<NR>(acc, partials, counter, lb, nb)) {
            double tot[NR];
            gather_partials<NR>(tot, partials, nb);
            if (threadIdx.x == 0) fin(tot);
        }
    }
}

struct NoFin {
    template <class T>
    __device__ void operator()(const T&) const {}
};

}  // namespace fcfd
