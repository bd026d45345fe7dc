// Passive scalar transport (jax_fem_cfd/solvers/passive_transport.py): Crank-Nicolson
// advection-diffusion of phi by BiCGSTAB, plain (update_phi :39-61) or SUPG-stabilised
// (update_phi_SUPG :14-37), after the Navier-Stokes step (timestep :64-97).
//
// One element kernel builds the two per-element matrices of the scalar system from U_n and U:
//   Mt  = Me + tau Ce_n^T                       (scalar_consistent_mass_SUPG_mult)
//   K*  = Ce* + D Le + tau Ct_n                 (scalar_stabilized_visc_conv_mult)
//   A_e = Mt / dt + (Ce(U)   + D Le + tau Ct_n) / 2     (lhs operator)
//   B_e = Mt / dt - (Ce(U_n) + D Le + tau Ct_n) / 2     (rhs: B phi_n)
// with Ce(., beta = 1) (passive_transport.py:80-81), Ct_n = Ctau_tilde(U_n)
// (element_calc.py:188-200) and tau from calculate_tau with r = 2 (element_calc.py:202-229;
// tau = 0 without SUPG).  The scalar BiCGSTAB is the generic one with one component.
#include <cmath>
#include "fcfd_bicgstab.hpp"

namespace fcfd {

// Ct[a][b] = sum_q w (U_q . grad N_a)(U_q . grad N_b)
__device__ __forceinline__ void ctau_elem(const ElemConst* __restrict__ ec, const double (&ul)[4],
                                          const double (&vl)[4], double (&ct)[16], double& u2mean) {
#pragma unroll
    for (int k = 0; k < 16; ++k) ct[k] = 0.0;
    u2mean = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double uq = 0.0, vq = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            uq += ec->N[q * 4 + c] * ul[c];
            vq += ec->N[q * 4 + c] * vl[c];
        }
        double gN[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) gN[a] = uq * ec->dNx[q * 4 + a] + vq * ec->dNy[q * 4 + a];
        const double w = ec->wq[q];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) ct[a * 4 + b] += w * gN[a] * gN[b];
        u2mean += uq * uq + vq * vq;
    }
    u2mean *= 0.25;
}

// tau = (tau1^-r + tau2^-r + tau3^-r)^(-1/r), tau1 = |Ce|/|Ct|, tau2 = dt/2, tau3 = tau1 Re,
// Re = mean|U_q|^2 |Ce| / ((D + 1e-6) |Ct|)   (Frobenius norms; element_calc.py:202-229)
__device__ __forceinline__ double tau_elem(const double (&ce)[16], const double (&ct)[16], double u2mean, double D,
                                           double dt, double r) {
    double nce = 0.0, nct = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        nce += ce[k] * ce[k];
        nct += ct[k] * ct[k];
    }
    nce = sqrt(nce);
    nct = sqrt(nct);
    const double Dp = D + 1e-6;
    const double Re = (u2mean * nce) / (Dp * nct);
    const double t1 = nce / nct, t2 = dt / 2, t3 = t1 * Re;
    if (r == 2.0) return pow(1.0 / (t1 * t1) + 1.0 / (t2 * t2) + 1.0 / (t3 * t3), -0.5);
    return pow(1.0 / pow(t1, r) + 1.0 / pow(t2, r) + 1.0 / pow(t3, r), -1.0 / r);
}

struct AdeParams {
    double dt, D;
    int supg;
};

__global__ __launch_bounds__(BLOCK) void k_ade_elem(Grid2D g, const ElemConst* __restrict__ ec,
                                                   const double* __restrict__ Un, const double* __restrict__ U,
                                                   AdeParams p, double* __restrict__ A, double* __restrict__ B) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (e >= g.ne) return;
    int cn[4];
    elem_nodes(g, e, cn);
    double un[4], vn[4], u[4], v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        un[c] = Un[cn[c]]; vn[c] = Un[cn[c] + g.nn];
        u[c] = U[cn[c]]; v[c] = U[cn[c] + g.nn];
    }
    double cen[16], ce[16], ct[16];
    conv_elem(ec, un, vn, 1.0, cen);
    conv_elem(ec, u, v, 1.0, ce);
    double tau = 0.0;
    if (p.supg) {
        double u2;
        ctau_elem(ec, un, vn, ct, u2);
        tau = tau_elem(ce, ct, u2, p.D, p.dt, 2.0);
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) ct[k] = 0.0;
    }
    const size_t ne = g.ne;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int k = a * 4 + b;
            const double Mt = ec->Me[k] + (p.supg ? tau * cen[b * 4 + a] : 0.0);
            const double K = p.D * ec->Le[k] + (p.supg ? tau * ct[k] : 0.0);
            A[ce_idx(k, e, (int)ne)] = (1.0 / p.dt) * Mt + 0.5 * (ce[k] + K);
            B[ce_idx(k, e, (int)ne)] = (1.0 / p.dt) * Mt - 0.5 * (cen[k] + K);
        }
}

// Ctau_tilde (ne,4,4) and tau (ne) in the reference layout, for the element_calc API
__global__ void k_supg_api(Grid2D g, const ElemConst* __restrict__ ec, const double* __restrict__ Un,
                           const double* __restrict__ Ce_api, double D, double dt, double r, double* Ct_api,
                           double* tau) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.ne) return;
    int cn[4];
    elem_nodes(g, e, cn);
    double un[4], vn[4];
    for (int c = 0; c < 4; ++c) { un[c] = Un[cn[c]]; vn[c] = Un[cn[c] + g.nn]; }
    double ct[16], u2;
    ctau_elem(ec, un, vn, ct, u2);
    if (Ct_api)
        for (int k = 0; k < 16; ++k) Ct_api[(size_t)e * 16 + k] = ct[k];
    if (tau && Ce_api) {
        double ce[16];
        for (int k = 0; k < 16; ++k) ce[k] = Ce_api[(size_t)e * 16 + k];
        tau[e] = tau_elem(ce, ct, u2, D, dt, r);
    }
}

void supg_api(Ctx& c, const double* U_n, const double* Ce_api, double D, double dt, double r, double* Ct, double* tau) {
    k_supg_api<<<blocks_for(c.g.ne), BLOCK, 0, c.stream>>>(c.g, c.ec_dev, U_n, Ce_api, D, dt, r, Ct, tau);
    FCFD_CHECK_LAUNCH();
}

// Reverse-mode product of the SUPG element terms.  With w = the masked scalar adjoint, the
// step's scalar residual contributes l_e = w^T (B_e phi_n - A_e phi) per element, i.e.
//   l_e = w^T Me d / dt - D w^T Le s / 2 - w^T Ce(U) phi / 2 - w^T Ce(U_n) phi_n / 2
//         + tau d^T Ce(U_n) w / dt - tau w^T Ct(U_n) s / 2,        d = phi_n - phi, s = phi_n + phi.
// The two plain convection terms are gathered at the nodes (conv_grad_node); this kernel
// writes, per element and local node I, the remaining U_n and U derivatives:
//   tau-explicit:  d/dU_n [tau d^T Ce_n w / dt - tau w^T Ct s / 2]
//   through tau:   dl/dtau = d^T Ce_n w / dt - w^T Ct s / 2, tau = S^(-1/2),
//                  S = (|Ct|/|Ce|)^2 + (2/dt)^2 + ((D+1e-6)|Ct|^2 / (u2 |Ce|^2))^2
//                  with |Ce| = |Ce(U)|, |Ct| = |Ct(U_n)|, u2 = mean_q |U_n,q|^2 (r = 2).
// Output layout gel[((v * 4 + I) * 2 + d) * ne + e], v = 0 for U_n, 1 for U.
__global__ __launch_bounds__(BLOCK) void k_supg_adj_elem(Grid2D g, const ElemConst* __restrict__ ec,
                                                        const double* __restrict__ Un,
                                                        const double* __restrict__ U,
                                                        const double* __restrict__ mu,
                                                        const uint8_t* __restrict__ mphi,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ phin, double D, double dt,
                                                        double* __restrict__ gel) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (e >= g.ne) return;
    int cn[4];
    elem_nodes(g, e, cn);
    double un[4], vn[4], u[4], v[4], w[4], dl[4], sm[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        un[c] = Un[cn[c]]; vn[c] = Un[cn[c] + g.nn];
        u[c] = U[cn[c]]; v[c] = U[cn[c] + g.nn];
        const double m = mu[cn[c]];
        w[c] = (mphi && mphi[cn[c]]) ? 0.0 : m;
        dl[c] = phin[cn[c]] - phi[cn[c]];
        sm[c] = phin[cn[c]] + phi[cn[c]];
    }
    double ce[16], cen[16], ct[16], u2;
    conv_elem(ec, u, v, 1.0, ce);
    conv_elem(ec, un, vn, 1.0, cen);
    ctau_elem(ec, un, vn, ct, u2);
    double nce = 0.0, nct = 0.0, pt = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int k = a * 4 + b;
            nce += ce[k] * ce[k];
            nct += ct[k] * ct[k];
            pt += dl[a] * cen[k] * w[b] / dt - 0.5 * w[a] * ct[k] * sm[b];
        }
    nce = sqrt(nce);
    nct = sqrt(nct);
    const double Dp = D + 1e-6;
    const double A1 = nct / nce, A3 = Dp * nct * nct / (u2 * nce * nce);
    const double S = A1 * A1 + (2.0 / dt) * (2.0 / dt) + A3 * A3;
    const double tau = 1.0 / sqrt(S);
    const double gS = -0.5 * tau * tau * tau * pt;
    const double k13 = 2.0 * A1 * A1 + 4.0 * A3 * A3;
    const double cG = gS * (-k13 / nce) / nce;   // d l / d Ce(U)_ab = cG Ce_ab
    const double cH = gS * (k13 / nct) / nct;    // d l / d Ct_ab    = cH Ct_ab
    const double gu2 = gS * (-2.0 * A3 * A3 / u2);
    double gn[4][2], gu[4][2];
#pragma unroll
    for (int I = 0; I < 4; ++I) gn[I][0] = gn[I][1] = gu[I][0] = gu[I][1] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double* Nq = ec->N + q * 4;
        const double* Xq = ec->dNx + q * 4;
        const double* Yq = ec->dNy + q * 4;
        const double wt = ec->wq[q];
        double uq = 0.0, vq = 0.0, Dq = 0.0, Wq = 0.0, Wx = 0.0, Wy = 0.0, Sx = 0.0, Sy = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            uq += Nq[c] * un[c]; vq += Nq[c] * vn[c];
            Dq += Nq[c] * dl[c]; Wq += Nq[c] * w[c];
            Wx += Xq[c] * w[c]; Wy += Yq[c] * w[c];
            Sx += Xq[c] * sm[c]; Sy += Yq[c] * sm[c];
        }
        const double aW = uq * Wx + vq * Wy, aS = uq * Sx + vq * Sy;
        // Ce(U) chain: sum_ab G_ab N_a dN_b and sum_ab G_ab N_a N_b;  Ct chain: sum_ab H_ab dN_a (U.grad N_b)
        double gNx = 0.0, gNy = 0.0, gNN = 0.0, hx = 0.0, hy = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            double cNa = 0.0, tb = 0.0, tx = 0.0, ty = 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                cNa += ce[a * 4 + b] * Nq[a];
                tb = ct[a * 4 + b];
                tx += tb * Xq[a];
                ty += tb * Yq[a];
            }
            gNx += cNa * Xq[b];
            gNy += cNa * Yq[b];
            gNN += cNa * Nq[b];
            const double ab = uq * Xq[b] + vq * Yq[b];
            hx += tx * ab;
            hy += ty * ab;
        }
#pragma unroll
        for (int I = 0; I < 4; ++I) {
            const double NI = Nq[I];
            // U: cG * sum_ab Ce_ab dCe_ab/dU_I
            gu[I][0] += cG * wt * (NI * gNx + Xq[I] * gNN);
            gu[I][1] += cG * wt * (NI * gNy + Yq[I] * gNN);
            // U_n: (tau/dt) d/dU_n [d^T Ce_n w]
            const double cx = (tau / dt) * wt * Dq * (NI * Wx + Xq[I] * Wq);
            const double cy = (tau / dt) * wt * Dq * (NI * Wy + Yq[I] * Wq);
            // U_n: -(tau/2) d/dU_n [w^T Ct s]
            const double tx = -0.5 * tau * wt * NI * (Wx * aS + Sx * aW);
            const double ty = -0.5 * tau * wt * NI * (Wy * aS + Sy * aW);
            // U_n: cH * sum_ab Ct_ab dCt_ab/dU_n,I  and  gu2 * d u2 / dU_n,I
            gn[I][0] += cx + tx + 2.0 * cH * wt * NI * hx + 0.5 * gu2 * NI * uq;
            gn[I][1] += cy + ty + 2.0 * cH * wt * NI * hy + 0.5 * gu2 * NI * vq;
        }
    }
    const size_t ne = g.ne;
#pragma unroll
    for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            gel[((size_t)(0 * 4 + I) * 2 + d) * ne + e] = gn[I][d];
            gel[((size_t)(1 * 4 + I) * 2 + d) * ne + e] = gu[I][d];
        }
}

void supg_adjoint_elem(Ctx& c, const double* U_n, const double* U, const double* mu, const double* phi,
                       const double* phi_n, double D, double dt, double* gel) {
    k_supg_adj_elem<<<blocks_for(c.g.ne), BLOCK, 0, c.stream>>>(c.g, c.ec_dev, U_n, U, mu, c.phi_mask, phi, phi_n,
                                                                D, dt, gel);
    FCFD_CHECK_LAUNCH();
}

static void ensure_ade(Ctx& c) {
    if (!c.ade_A) {
        c.ade_A = dalloc<double>(c, 16 * (size_t)c.g.ne);
        c.ade_B = dalloc<double>(c, 16 * (size_t)c.g.ne);
        c.s_x = dalloc<double>(c, c.g.nn);
    }
}

__global__ void k_scatter_phi(int n, const int64_t* __restrict__ gd, const double* __restrict__ v, double* full) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) full[gd[k]] = v ? v[k] : 0.0;
}

int solve_scalar(Ctx& c, const double* U_n, const double* U, const double* phi_n, double dt, double D,
                 const double* phi_d, int supg, double* phi) {
    ensure_ade(c);
    k_ade_elem<<<blocks_for(c.g.ne), BLOCK, 0, c.stream>>>(c.g, c.ec_dev, U_n, U, AdeParams{dt, D, supg}, c.ade_A,
                                                           c.ade_B);
    FCFD_CHECK_LAUNCH();
    if (c.n_gamma_d_phi > 0) {
        k_scatter_phi<<<(c.n_gamma_d_phi + 255) / 256, 256, 0, c.stream>>>(c.n_gamma_d_phi, c.gamma_d_phi_dev, phi_d,
                                                                            c.phi_ud);
        FCFD_CHECK_LAUNCH();
    }
    StoredCoef<CeF64> lhs{c.g, CeF64{c.ade_A, c.g.ne}};
    StoredCoef<CeF64> rhs{c.g, CeF64{c.ade_B, c.g.ne}};
    const double key[4] = {0.0, 0.0, 0.0, 0.0};
    const int maxit = 10 * c.g.nn;
    const int it = bicgstab_solve<1>(c, c.g_phi, key, c.ade_A, lhs, rhs, NoExtra{}, phi_n, c.phi_ud, c.phi_mask,
                                     maxit, c.s_x, c.pred_phi);
    if (phi != c.s_x)
        FCFD_HIP(hipMemcpyAsync(phi, c.s_x, sizeof(double) * c.g.nn, hipMemcpyDeviceToDevice, c.stream));
    return it;
}

}  // namespace fcfd
