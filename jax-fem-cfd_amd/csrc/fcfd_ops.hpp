// Node-owner element-by-element operators.
//
// Every global matvec y = sum_e P_e^T A_e P_e x is evaluated per node (owner computes): a node
// gathers its 3x3 neighbourhood once and sums, in the fixed order of its incident elements
// (lower-left, lower-right, upper-left, upper-right), the row of each element matrix that
// belongs to it.  This is the scatter-add of the reference (jax.ops.segment_sum over nconn,
// jax_fem_cfd/core/element_mult.py) turned into a deterministic gather: one pass, no colouring,
// no float atomics, every per-element matrix entry read exactly once.
#pragma once
#include "fcfd_common.hpp"

namespace fcfd {

// 3x3 gather: f is evaluated at every (clamped, valid) neighbour id, so all loads issue
// back to back; missing neighbours are zeroed with a select afterwards.
template <class T, class F>
__device__ __forceinline__ void load3x3(const NodeNbr& s, F f, T (&X)[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const T v = f(s.nb[a][b]);
            X[a][b] = s.ok[a][b] ? v : T(0);
        }
}

// y(node) = sum_k sum_b A[row_k][b] X[corner_k(b)], one shared element matrix
template <class T>
__device__ __forceinline__ T ebe_const(const NodeNbr& s, const double* __restrict__ A, const T (&X)[3][3]) {
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = kRow(k);
        T t = T(0);
#pragma unroll
        for (int b = 0; b < 4; ++b) t += T(A[a * 4 + b]) * X[kCy(k, b)][kCx(k, b)];
        acc += s.has[k] ? t : T(0);
    }
    return acc;
}

// same with A^T (divergence uses G_i^T)
template <class T>
__device__ __forceinline__ T ebe_const_T(const NodeNbr& s, const double* __restrict__ A, const T (&X)[3][3]) {
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = kRow(k);
        T t = T(0);
#pragma unroll
        for (int b = 0; b < 4; ++b) t += T(A[b * 4 + a]) * X[kCy(k, b)][kCx(k, b)];
        acc += s.has[k] ? t : T(0);
    }
    return acc;
}

// Element corner node ids
__device__ __forceinline__ void elem_nodes(const Grid2D& g, int e, int (&c)[4]) {
    int je = e / g.nex, ie = e - je * g.nex;
    if (g.periodic) {
        int ip = ie + 1 == g.nnx ? 0 : ie + 1;
        int jp = je + 1 == g.nny ? 0 : je + 1;
        c[0] = je * g.nnx + ie; c[1] = je * g.nnx + ip;
        c[2] = jp * g.nnx + ip; c[3] = jp * g.nnx + ie;
    } else {
        int n0 = je * g.nnx + ie;
        c[0] = n0; c[1] = n0 + 1; c[2] = n0 + g.nnx + 1; c[3] = n0 + g.nnx;
    }
}

// Per-element convection matrix Ce = Ce_A + beta Ce_D from the element's nodal velocities
// (jax_fem_cfd/core/element_calc.py:165-185):
//   Ce_A[a][b] = sum_q w_q N_qa (U_q . grad N_qb),   U_q = sum_c N_qc U_c
//   Ce_D[a][b] = sum_q w_q divU_q N_qa N_qb,          divU_q = sum_d sum_c d_d N_qc U_cd
__device__ __forceinline__ void conv_elem(const ElemConst* __restrict__ ec, const double (&ul)[4],
                                          const double (&vl)[4], double beta, double (&ce)[16]) {
    double A[16], D[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) { A[k] = 0.0; D[k] = 0.0; }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double uq = 0.0, vq = 0.0, dx = 0.0, dy = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            uq += ec->N[q * 4 + c] * ul[c];
            vq += ec->N[q * 4 + c] * vl[c];
            dx += ec->dNx[q * 4 + c] * ul[c];
            dy += ec->dNy[q * 4 + c] * vl[c];
        }
        const double div = dx + dy;
        const double w = ec->wq[q];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double wNa = ec->N[q * 4 + a] * w;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                double adv = uq * ec->dNx[q * 4 + b] + vq * ec->dNy[q * 4 + b];
                A[a * 4 + b] += wNa * adv;
                D[a * 4 + b] += (ec->N[q * 4 + a] * w * div) * ec->N[q * 4 + b];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) ce[k] = A[k] + beta * D[k];
}

}  // namespace fcfd
