// Shared device infrastructure: structured-grid index math, deterministic
// two-stage reductions, XCD-aware block mapping, element-matrix storage.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <type_traits>
#include <cstdio>
#include <string>

namespace fcfd {

constexpr int BLOCK = 256;         // threads per node/element kernel block (4 waves)
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

// The gate flag of a kernel enqueued ahead of its solver's convergence test (null: always
// open), read at entry but tested late: a test at entry would put the flag's memory round trip in
// front of every other load of the block.  The kernels that use this test it once the block's
// first loads are issued -- after their first barrier, or before their first global store --
// and until then only read, at valid (clamped) addresses.  The value is the same for the whole
// grid, so the test is block-uniform and may sit between barriers.  The flag is read with a
// vector-memory load: scalar loads share one wait counter with the kernel argument loads, so the
// first use of any argument would wait for the flag too.  gate_read issues the load -- in most
// kernels right behind the block's first batch of loads: issued ahead of them, it makes the block
// wait for the flag pointer's kernel-argument load before it can issue any other load (1.4% of a
// 1024^2 cavity step, measured) -- and gate_open tests the value where the block needs it.  The load is a raw buffer load with the
// agent-scope coherence bit (sc1) -- what an agent-scope atomic load compiles to, since other
// kernels on other CUs and XCDs write the flag -- issued through the builtin rather than as an
// atomic, which would count as a write and push the kernel's uniform loads of constant data
// (element matrices) off the scalar path as well.  A null gate reads an always-open flag
// instead of branching.  The finalize kernels, which write the solver state the flag lives in,
// test it at entry instead: with the late test, the 1-block finalize of alpha in a hex27 PCG read
// the flag as still closed for the first iterations of a solve on one box (3 of 3 runs of the
// GPU suite; the same build with the entry test passed there).
namespace {
__device__ int kGateAlwaysOpen = 1;
}
// Buffer resource word 3 for a raw 32-bit buffer load on CDNA (DATA_FORMAT = 32 bits), and the
// cache-policy operand with sc1 set (agent-scope coherence on gfx94x / gfx950)
constexpr int kBufRsrcWord3 = 0x00020000;
constexpr int kCachePolicySc1 = 16;
__device__ __forceinline__ int gate_read(const int* active) {
    const int* p = active ? active : &kGateAlwaysOpen;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, 4, kBufRsrcWord3);
    return __builtin_amdgcn_raw_buffer_load_b32(rs, 0, 0, kCachePolicySc1);
}
// the test, at the site where the block needs it
__device__ __forceinline__ bool gate_open(int g) { return __builtin_amdgcn_readfirstlane(g) != 0; }

// The solver-iteration kernels with the late test also come as EARLY instantiations, which test
// the flag at entry too, before any other load (the finalize kernels' scalar load; the flag is
// then non-null).  The host launches them for the iterations of a speculative chunk past the
// count it is confident of: those mostly find the gate closed and then end without moving the
// block's inputs, which the late test cannot spare them (their loads are in flight when it
// runs).  A compile-time choice, so that the late-test launches keep their single batch of
// kernel-argument loads.

// Stencil context of one node: the four incident elements (element to the lower-left,
// lower-right, upper-left, upper-right of the node) with the node's local index in each
// (2, 3, 1, 0), and the 3x3 neighbour node ids.  Ids of missing neighbours / elements (edges of a
// non-periodic grid) are clamped to valid ones (the node itself, element 0) and flagged, so every
// load can be issued unconditionally and masked afterwards with a select -- no per-lane branches
// around memory accesses.
struct NodeNbr {
    int i, j;
    int e[4];       // element ids (clamped to 0 where !has)
    bool has[4];
    int nb[3][3];   // nb[dy+1][dx+1] node id (clamped to the node itself where !ok)
    bool ok[3][3];
};

__device__ __forceinline__ void node_nbr(const Grid2D& g, int n, NodeNbr& s) {
    int j = n / g.nnx, i = n - j * g.nnx;
    s.i = i; s.j = j;
    if (g.periodic) {
        int im = i == 0 ? g.nnx - 1 : i - 1, ip = i == g.nnx - 1 ? 0 : i + 1;
        int jm = j == 0 ? g.nny - 1 : j - 1, jp = j == g.nny - 1 ? 0 : j + 1;
        int xs[3] = {im, i, ip}, ys[3] = {jm, j, jp};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                s.nb[a][b] = ys[a] * g.nnx + xs[b];
                s.ok[a][b] = true;
            }
        s.e[0] = jm * g.nex + im; s.e[1] = jm * g.nex + i;
        s.e[2] = j * g.nex + im;  s.e[3] = j * g.nex + i;
        s.has[0] = s.has[1] = s.has[2] = s.has[3] = true;
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int jj = j + a - 1;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int ii = i + b - 1;
                const bool ok = (ii >= 0) && (ii < g.nnx) && (jj >= 0) && (jj < g.nny);
                s.ok[a][b] = ok;
                s.nb[a][b] = ok ? jj * g.nnx + ii : n;
            }
        }
        const bool l = i >= 1, r = i < g.nex, d = j >= 1, u = j < g.ney;
        s.has[0] = l && d; s.has[1] = r && d; s.has[2] = l && u; s.has[3] = r && u;
        s.e[0] = s.has[0] ? (j - 1) * g.nex + (i - 1) : 0;
        s.e[1] = s.has[1] ? (j - 1) * g.nex + i : 0;
        s.e[2] = s.has[2] ? j * g.nex + (i - 1) : 0;
        s.e[3] = s.has[3] ? j * g.nex + i : 0;
    }
}

// Full-wavefront sum with DPP lane moves (quad_perm, row_shr, row_bcast; gfx9 family): the
// total lands in lane 63 and is broadcast with readlane.  FP64 moves go as two 32-bit halves.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
    const int l2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, false);
    const int h2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double(((long long)h2 << 32) | (unsigned int)l2);
}

__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_move<0xb1>(v);        // quad_perm [1,0,3,2]
    v += dpp_move<0x4e>(v);        // quad_perm [2,3,0,1]
    v += dpp_move<0x114>(v);       // row_shr:4
    v += dpp_move<0x118>(v);       // row_shr:8
    v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
    v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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

// quad9 (Q2) element constants: 3x3 Gauss data [q][I] and the 9 x 9 matrices
struct ElemConstQ2 {
    double N[81], dNx[81], dNy[81], wq[9];
    double Me[81], Le[81], Leoff[81], Gx[81], Gy[81];
};

// ---------------------------------------------------------------------------------------------
// Per-element matrix storage, row-blocked: entry (a, b) of element e at ((a * ne + e) * 4 + b),
// i.e. the 4 entries of row a of an element are contiguous and row a of consecutive elements is
// consecutive.  A node reads the row of each incident element that belongs to it with one vector
// load (16 B FP32, 32 B FP64, 8 B FP16 / BF16), and a wavefront reading row a of 64 consecutive
// elements reads one contiguous span (coalesced).
// ---------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ size_t ce_idx(int k, int e, int ne) {
    return ((size_t)(k >> 2) * ne + e) * 4 + (k & 3);
}

struct CeF64 {
    double* c; int ne;
    __device__ __forceinline__ double ld(int k, int e) const { return c[ce_idx(k, e, ne)]; }
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[4]) const {
        const double2* p = reinterpret_cast<const double2*>(c + ce_idx(a * 4, e, ne));
        const double2 x = p[0], y = p[1];
        v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
    }
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        for (int k = 0; k < 16; ++k) c[ce_idx(k, e, ne)] = v[k];
    }
    static constexpr size_t bytes_per_elem = 16 * 8;
};
struct CeF32 {
    float* c; int ne;
    __device__ __forceinline__ double ld(int k, int e) const { return (double)c[ce_idx(k, e, ne)]; }
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[4]) const {
        const float4 x = *reinterpret_cast<const float4*>(c + ce_idx(a * 4, e, ne));
        v[0] = (double)x.x; v[1] = (double)x.y; v[2] = (double)x.z; v[3] = (double)x.w;
    }
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        for (int k = 0; k < 16; ++k) c[ce_idx(k, e, ne)] = (float)v[k];
    }
    static constexpr size_t bytes_per_elem = 16 * 4;
};
// FP16 entries normalised by a per-element FP32 scale (max |entry|), 36 B / element
struct CeF16S {
    __half* c; float* s; int ne;
    __device__ __forceinline__ double ld(int k, int e) const {
        return (double)(__half2float(c[ce_idx(k, e, ne)]) * s[e]);
    }
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[4]) const {
        const uint2 w = *reinterpret_cast<const uint2*>(c + ce_idx(a * 4, e, ne));
        const float sc = s[e];
        const __half2 h0 = *reinterpret_cast<const __half2*>(&w.x), h1 = *reinterpret_cast<const __half2*>(&w.y);
        v[0] = (double)(__low2float(h0) * sc); v[1] = (double)(__high2float(h0) * sc);
        v[2] = (double)(__low2float(h1) * sc); v[3] = (double)(__high2float(h1) * sc);
    }
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        double m = 0.0;
        for (int k = 0; k < 16; ++k) m = fmax(m, fabs(v[k]));
        float sc = m > 0.0 ? (float)m : 1.0f;
        s[e] = sc;
        for (int k = 0; k < 16; ++k) c[ce_idx(k, e, ne)] = __float2half((float)(v[k] / (double)sc));
    }
    static constexpr size_t bytes_per_elem = 16 * 2 + 4;
};
struct CeBF16S {
    __hip_bfloat16* c; float* s; int ne;
    __device__ __forceinline__ double ld(int k, int e) const {
        return (double)(__bfloat162float(c[ce_idx(k, e, ne)]) * s[e]);
    }
    __device__ __forceinline__ void ld_row(int a, int e, double (&v)[4]) const {
        const uint2 w = *reinterpret_cast<const uint2*>(c + ce_idx(a * 4, e, ne));
        const float sc = s[e];
        // bf16 -> f32: the 16 bits become the high half of the float
        v[0] = (double)(__uint_as_float(w.x << 16) * sc); v[1] = (double)(__uint_as_float(w.x & 0xffff0000u) * sc);
        v[2] = (double)(__uint_as_float(w.y << 16) * sc); v[3] = (double)(__uint_as_float(w.y & 0xffff0000u) * sc);
    }
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        double m = 0.0;
        for (int k = 0; k < 16; ++k) m = fmax(m, fabs(v[k]));
        float sc = m > 0.0 ? (float)m : 1.0f;
        s[e] = sc;
        for (int k = 0; k < 16; ++k) c[ce_idx(k, e, ne)] = __float2bfloat16((float)(v[k] / (double)sc));
    }
    static constexpr size_t bytes_per_elem = 16 * 2 + 4;
};

// ---------------------------------------------------------------------------------------------
// Deterministic reductions.  Each reducing kernel accumulates NR per-thread sums, reduces them
// per block in a fixed tree, writes one partial per block, and the last block to finish sums
// the partials in fixed order and runs the kernel's finalize step (device scalar updates).
// No float atomics anywhere: results are bitwise reproducible run to run.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) { return wave_sum_dpp(v); }

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

// One item (node / element) per thread; logical block lb owns items [lb*BLOCK, lb*BLOCK+BLOCK).
inline int blocks_for(int n) { return n > 0 ? (n + BLOCK - 1) / BLOCK : 1; }

constexpr int FIN_BLOCK = 512;   // threads of the one-block finalize kernel (measured: 512 > 1024 > 256)

struct NoFin {
    template <class T>
    __device__ void operator()(const T&) const {}
};

// Epilogue of a reducing kernel: the block totals go to partials[r * pstride + lb]; the
// finalize step runs as its own one-block kernel (finalize_kernel) right after.
template <int NR>
__device__ __forceinline__ void reduce_epilogue(double (&acc)[NR], double* __restrict__ partials, int pstride,
                                                int lb) {
    block_sum<NR>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) partials[r * pstride + lb] = acc[r];
    }
}

// Generic node kernel with NR fused reductions.  Body: void operator()(int n, double (&acc)[NR]).
// Each block reduces its per-thread sums in a fixed tree and stores one partial per reduction
// (plain stores); finalize_kernel (one block, launched right after on the same stream) sums the
// partials in fixed order and runs the solver's scalar update.  No atomics and no grid-scope
// fences: visibility comes from the kernel boundary, and results are bitwise reproducible.
// `active` (device flag, may be null) lets a whole chunk of solver iterations be enqueued ahead:
// once the solver has converged every later kernel returns at entry.
// BS: threads per block (a multiple of 64).
template <int NR, class Body, int BS = BLOCK>
__global__ __launch_bounds__(BS) void node_kernel(int n, const int* __restrict__ active, Body body,
                                                 double* __restrict__ partials, int pstride) {
    if (active && *active == 0) return;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int k = lb * BS + threadIdx.x;
    double acc[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < (NR > 0 ? NR : 1); ++r) acc[r] = 0.0;
    if (k < n) body(k, acc);
    if constexpr (NR > 0) reduce_epilogue<NR>(acc, partials, pstride, lb);
}

// ---------------------------------------------------------------------------------------------
// LDS-staged stencil kernels (non-periodic grids).  The 3x3 neighbours of the nodes
// [n0, n0 + BLOCK) of a block are the linear ids n + dy*nnx + dx, which all lie in three
// contiguous windows [n0 + dy*nnx - 1, n0 + dy*nnx + BLOCK + 1).  The block stages those windows
// once (coalesced loads; the body computes the staged value, e.g. r - alpha q, once per node
// instead of once per stencil point) and every node reads its stencil from LDS.  Ids that wrap
// across a row end belong only to missing elements (masked by has[k]); ids outside [0, nn) are
// staged as 0.  Per node: ~3 global loads per staged vector instead of 9.
// ---------------------------------------------------------------------------------------------
// Window width for NPT nodes per thread (a block owns BLOCK * NPT consecutive nodes).
template <int NPT>
struct SWin {
    static constexpr int W = BLOCK * NPT + 2;
};

template <int NV, int NPT = 1>
struct StageView {
    const double (*s)[3][SWin<NPT>::W];
    int n0, nnx;
    __device__ __forceinline__ double at(int v, int n, int dy, int dx) const { return s[v][dy + 1][n - n0 + dx + 1]; }
};

// Incident elements of node n of a non-periodic grid (ids clamped to 0 where missing).
__device__ __forceinline__ void node_elems(const Grid2D& g, int n, NodeNbr& s) {
    const int j = n / g.nnx, i = n - j * g.nnx;
    s.i = i; s.j = j;
    const bool l = i >= 1, r = i < g.nex, d = j >= 1, u = j < g.ney;
    s.has[0] = l && d; s.has[1] = r && d; s.has[2] = l && u; s.has[3] = r && u;
    s.e[0] = s.has[0] ? (j - 1) * g.nex + (i - 1) : 0;
    s.e[1] = s.has[1] ? (j - 1) * g.nex + i : 0;
    s.e[2] = s.has[2] ? j * g.nex + (i - 1) : 0;
    s.e[3] = s.has[3] ? j * g.nex + i : 0;
}

// Body: static constexpr int NV; static constexpr int NPT; int nnx() const;
//       void stage(int id, double (&v)[NV]) const;
//       void operator()(int n, const StageView<NV, NPT>&, double (&acc)[NR]) const.
// Staging order.  By default a thread issues all of its staging loads before its first LDS store
// (their latencies overlap: the short stencil kernels gain).  A body that streams much more than
// it stages (the momentum operator apply, whose element matrices follow the staging) sets
// LAZY_STAGE: one load-store pair at a time measured faster there.
template <class B, class = void>
struct LazyStage {
    static constexpr bool value = false;
};
template <class B>
struct LazyStage<B, std::void_t<decltype(B::LAZY_STAGE)>> {
    static constexpr bool value = B::LAZY_STAGE;
};

// A block owns BLOCK * NPT consecutive nodes; thread t computes nodes n0 + t + m * BLOCK.
template <int NR, class Body>
__global__ __launch_bounds__(BLOCK) void stencil_kernel(int n, const int* __restrict__ active, Body body,
                                                       double* __restrict__ partials, int pstride) {
    if (active && *active == 0) return;
    constexpr int NV = Body::NV, NPT = Body::NPT, W = SWin<NPT>::W;
    __shared__ double sm[NV][3][W];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int n0 = lb * BLOCK * NPT;
    const int nnx = body.nnx();
    if constexpr (LazyStage<Body>::value) {
        for (int idx = threadIdx.x; idx < 3 * W; idx += BLOCK) {
            const int seg = idx / W, off = idx - seg * W;
            const int id = n0 + (seg - 1) * nnx - 1 + off;
            const bool in = id >= 0 && id < n;
            double v[NV];
            body.stage(in ? id : 0, v);
#pragma unroll
            for (int q = 0; q < NV; ++q) sm[q][seg][off] = in ? v[q] : 0.0;
        }
    } else {
        // a fixed trip count, unrolled: all loads in flight before the first LDS store
        constexpr int NIT = (3 * W + BLOCK - 1) / BLOCK;
        double v[NIT][NV];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = threadIdx.x + it * BLOCK;
            const int seg = idx / W, off = idx - seg * W;
            const int id = n0 + (seg - 1) * nnx - 1 + off;
            const bool in = idx < 3 * W && id >= 0 && id < n;
            body.stage(in ? id : 0, v[it]);
#pragma unroll
            for (int q = 0; q < NV; ++q) v[it][q] = in ? v[it][q] : 0.0;
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = threadIdx.x + it * BLOCK;
            if (idx < 3 * W) {
                const int seg = idx / W, off = idx - seg * W;
#pragma unroll
                for (int q = 0; q < NV; ++q) sm[q][seg][off] = v[it][q];
            }
        }
    }
    __syncthreads();
    double acc[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < (NR > 0 ? NR : 1); ++r) acc[r] = 0.0;
#pragma unroll
    for (int m = 0; m < NPT; ++m) {
        const int k = n0 + threadIdx.x + m * BLOCK;
        if (k < n) body(k, StageView<NV, NPT>{sm, n0, nnx}, acc);
    }
    if constexpr (NR > 0) reduce_epilogue<NR>(acc, partials, pstride, lb);
}

// Fin: void operator()(const double (&tot)[NR]) run by thread 0 with the grid totals.
template <int NR, class Fin>
__global__ __launch_bounds__(FIN_BLOCK) void finalize_kernel(int nb, const int* __restrict__ active, Fin fin,
                                                            const double* __restrict__ partials, int pstride) {
    if (active && *active == 0) return;
    // The partials were written by blocks on every XCD, so each read is a far-memory round trip:
    // a thread issues FIN_U loads per reduction before it adds any of them (fixed order, so the
    // sums stay bitwise reproducible), and 4096 partials take one round trip instead of four.
    constexpr int FIN_U = 8;
    double tot[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) tot[r] = 0.0;
    for (int k0 = threadIdx.x; k0 < nb; k0 += FIN_BLOCK * FIN_U) {
        double v[NR][FIN_U];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int u = 0; u < FIN_U; ++u) {
                const int k = k0 + u * FIN_BLOCK;
                v[r][u] = k < nb ? partials[r * pstride + k] : 0.0;
            }
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int u = 0; u < FIN_U; ++u) tot[r] += v[r][u];
    }
    block_sum<NR>(tot);
    if (threadIdx.x == 0) fin(tot);
}

// Two groups of partials with different block counts: slots [0, NR1) hold nb1 partials each,
// slots [NR1, NR1 + NR2) hold nb2 each (two reducing kernels feeding one scalar update).  Each
// slot is summed in the fixed order of finalize_kernel.
template <int NR1, int NR2, class Fin>
__global__ __launch_bounds__(FIN_BLOCK) void finalize2_kernel(int nb1, int nb2, const int* __restrict__ active, Fin fin,
                                                             const double* __restrict__ partials, int pstride) {
    if (active && *active == 0) return;
    constexpr int NR = NR1 + NR2, FIN_U = 8;
    double tot[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) tot[r] = 0.0;
    const int nbm = nb1 > nb2 ? nb1 : nb2;
    for (int k0 = threadIdx.x; k0 < nbm; k0 += FIN_BLOCK * FIN_U) {
        double v[NR][FIN_U];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int u = 0; u < FIN_U; ++u) {
                const int k = k0 + u * FIN_BLOCK;
                v[r][u] = k < (r < NR1 ? nb1 : nb2) ? partials[r * pstride + k] : 0.0;
            }
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int u = 0; u < FIN_U; ++u) tot[r] += v[r][u];
    }
    block_sum<NR>(tot);
    if (threadIdx.x == 0) fin(tot);
}

}  // namespace fcfd
