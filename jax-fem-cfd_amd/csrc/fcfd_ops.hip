// Element constants, convection-matrix build, element-by-element matvecs, boundary integral,
// mesh fill and BLAS-1 helpers.
#include <cmath>
#include <cstring>
#include "fcfd_internal.hpp"
#include "fcfd_ops.hpp"

namespace fcfd {

// ---------------------------------------------------------------------------------------------
// Host: quad4 2x2-Gauss data on element 0 of a uniform grid and the shared element matrices
// (jax_fem_cfd/core/shape_functions.py:27-45, core/element_calc.py:9-40).
// ---------------------------------------------------------------------------------------------
void quad4_constants(double hx, double hy, ElemConst& ec) {
    const double s = 1.0 / std::sqrt(3.0);
    const double xi[4] = {-s, s, s, -s}, eta[4] = {-s, -s, s, s};
    const double X[4] = {0.0, hx, hx, 0.0}, Y[4] = {0.0, 0.0, hy, hy};
    for (int q = 0; q < 4; ++q) {
        double N[4] = {(1 - xi[q]) * (1 - eta[q]) / 4, (1 + xi[q]) * (1 - eta[q]) / 4,
                       (1 + xi[q]) * (1 + eta[q]) / 4, (1 - xi[q]) * (1 + eta[q]) / 4};
        double dxi[4] = {-(1 - eta[q]) / 4, (1 - eta[q]) / 4, (1 + eta[q]) / 4, -(1 + eta[q]) / 4};
        double deta[4] = {-(1 - xi[q]) / 4, -(1 + xi[q]) / 4, (1 + xi[q]) / 4, (1 - xi[q]) / 4};
        double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
        for (int I = 0; I < 4; ++I) {
            J00 += dxi[I] * X[I]; J01 += dxi[I] * Y[I];
            J10 += deta[I] * X[I]; J11 += deta[I] * Y[I];
        }
        double det = J00 * J11 - J01 * J10;
        double i00 = J11 / det, i01 = -J01 / det, i10 = -J10 / det, i11 = J00 / det;
        ec.wq[q] = det;
        for (int I = 0; I < 4; ++I) {
            ec.N[q * 4 + I] = N[I];
            ec.dNx[q * 4 + I] = i00 * dxi[I] + i01 * deta[I];
            ec.dNy[q * 4 + I] = i10 * dxi[I] + i11 * deta[I];
        }
    }
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            double me = 0, le = 0, gx = 0, gy = 0;
            for (int q = 0; q < 4; ++q) {
                double w = ec.wq[q];
                me += w * ec.N[q * 4 + a] * ec.N[q * 4 + b];
                le += w * (ec.dNx[q * 4 + a] * ec.dNx[q * 4 + b] + ec.dNy[q * 4 + a] * ec.dNy[q * 4 + b]);
                gx += w * ec.N[q * 4 + a] * ec.dNx[q * 4 + b];
                gy += w * ec.N[q * 4 + a] * ec.dNy[q * 4 + b];
            }
            ec.Me[a * 4 + b] = me; ec.Le[a * 4 + b] = le;
            ec.Gx[a * 4 + b] = gx; ec.Gy[a * 4 + b] = gy;
            ec.Leoff[a * 4 + b] = (a == b) ? 0.0 : le;
        }
}

// ---------------------------------------------------------------------------------------------
// fills
// ---------------------------------------------------------------------------------------------
template <class T>
__global__ void k_fill(T* x, T v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
void launch_fill(Ctx& c, double* x, double v, int n) {
    k_fill<double><<<(n + 255) / 256, 256, 0, c.stream>>>(x, v, n);
    FCFD_CHECK_LAUNCH();
}
void launch_fill32(Ctx& c, float* x, float v, int n) {
    k_fill<float><<<(n + 255) / 256, 256, 0, c.stream>>>(x, v, n);
    FCFD_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------------------
// Ce build: one thread per element, row-blocked store (ce_idx).
// ---------------------------------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(BLOCK) void k_build_ce(Grid2D g, const ElemConst* __restrict__ ec,
                                                   const double* __restrict__ U, double beta, S store,
                                                   const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int e = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (e < g.ne) {
        int cn[4];
        elem_nodes(g, e, cn);
        double ul[4], vl[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) { ul[c] = U[cn[c]]; vl[c] = U[cn[c] + g.nn]; }
        double ce[16];
        conv_elem(ec, ul, vl, beta, ce);
        store.st_elem(e, ce);
    }
}

void build_ce(Ctx& c, const double* U, double beta, int fmt, void* store, float* scale) {
    int nb = blocks_for(c.g.ne);
    switch (fmt) {
        case FCFD_CE_F64:
            k_build_ce<<<nb, BLOCK, 0, c.stream>>>(c.g, c.ec_dev, U, beta, CeF64{(double*)store, c.g.ne}, c.gate);
            break;
        case FCFD_CE_F32:
            k_build_ce<<<nb, BLOCK, 0, c.stream>>>(c.g, c.ec_dev, U, beta, CeF32{(float*)store, c.g.ne}, c.gate);
            break;
        case FCFD_CE_F16S:
            k_build_ce<<<nb, BLOCK, 0, c.stream>>>(c.g, c.ec_dev, U, beta, CeF16S{(__half*)store, scale, c.g.ne}, c.gate);
            break;
        case FCFD_CE_BF16S:
            k_build_ce<<<nb, BLOCK, 0, c.stream>>>(c.g, c.ec_dev, U, beta,
                                                   CeBF16S{(__hip_bfloat16*)store, scale, c.g.ne}, c.gate);
            break;
        default:
            throw Error(FCFD_ERR_ARG, "unknown Ce format");
    }
    FCFD_CHECK_LAUNCH();
}

// Ce straight into the reference layout (ne,4,4)
struct CeAPIStore {
    double* c;
    __device__ __forceinline__ void st_elem(int e, const double* v) const {
        for (int k = 0; k < 16; ++k) c[(size_t)e * 16 + k] = v[k];
    }
};
void build_ce_api(Ctx& c, const double* U, double beta, double* Ce) {
    k_build_ce<<<blocks_for(c.g.ne), BLOCK, 0, c.stream>>>(c.g, c.ec_dev, U, beta, CeAPIStore{Ce}, nullptr);
    FCFD_CHECK_LAUNCH();
}

// (ne,4,4) row-major  <->  row-blocked (ce_idx)
__global__ void k_ce_layout(const double* __restrict__ in, double* __restrict__ out, int ne, int to_api) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    for (int k = 0; k < 16; ++k) {
        if (to_api) out[(size_t)e * 16 + k] = in[ce_idx(k, e, ne)];
        else out[ce_idx(k, e, ne)] = in[(size_t)e * 16 + k];
    }
}
void ce_to_api(Ctx& c, const double* em, double* api) {
    k_ce_layout<<<(c.g.ne + 255) / 256, 256, 0, c.stream>>>(em, api, c.g.ne, 1);
    FCFD_CHECK_LAUNCH();
}
void ce_from_api(Ctx& c, const double* api, double* em) {
    k_ce_layout<<<(c.g.ne + 255) / 256, 256, 0, c.stream>>>(api, em, c.g.ne, 0);
    FCFD_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------------------
// Shared-matrix EBE product y (+)= A x on grid g (element_mult.py:11-77, :133-157)
// ---------------------------------------------------------------------------------------------
struct ConstMat {
    double A[16];
};

__global__ __launch_bounds__(BLOCK) void k_const_apply(Grid2D g, ConstMat m, const double* __restrict__ x,
                                                      double* __restrict__ y, int acc) {
    const int n = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (n < g.nn) {
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3];
        load3x3(s, [&](int k) { return x[k]; }, X);
        double v = ebe_const(s, m.A, X);
        y[n] = acc ? y[n] + v : v;
    }
}

void elem_mult_const(Ctx& c, int level, const double* A16, const double* x, double* y, int acc) {
    FCFD_REQUIRE(level >= 0 && level < (int)c.lv.size(), "level out of range");
    ConstMat m;
    for (int k = 0; k < 16; ++k) m.A[k] = A16[k];
    const Grid2D& g = c.lv[level].g;
    k_const_apply<<<blocks_for(g.nn), BLOCK, 0, c.stream>>>(g, m, x, y, acc);
    FCFD_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------------------
// Stored-matrix products in the reference layout (ne,4,4):
//   y_c = sum_e (K0 + c1 A1_e + c2 A2_e) x_c(e),  c = 0..ncomp-1
// covers visc_conv_mult (element_mult.py:79-104: K0 = mu Le, A1 = Ce, c1 = rho),
// scalar_visc_conv_mult (:159-172), scalar_stabilized_visc_conv_mult (:174-187: A2 = Se) and
// scalar_consistent_mass_SUPG_mult (:190-201: K0 = Me, A1 = Me_tau).
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_stored_general(Grid2D g, ConstMat K0, const double* __restrict__ A1,
                                                         double c1, const double* __restrict__ A2, double c2,
                                                         int ncomp, const double* __restrict__ x,
                                                         double* __restrict__ y) {
    const int n = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (n < g.nn) {
        NodeNbr s;
        node_nbr(g, n, s);
        for (int comp = 0; comp < ncomp; ++comp) {
            const double* xc = x + (size_t)comp * g.nn;
            double X[3][3];
            load3x3(s, [&](int k) { return xc[k]; }, X);
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int a = kRow(k);
                const size_t eo = (size_t)s.e[k] * 16 + a * 4;
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    double coef = K0.A[a * 4 + b];
                    if (A1) coef += c1 * A1[eo + b];
                    if (A2) coef += c2 * A2[eo + b];
                    t += coef * X[kCy(k, b)][kCx(k, b)];
                }
                acc += s.has[k] ? t : 0.0;
            }
            y[(size_t)comp * g.nn + n] = acc;
        }
    }
}

void elem_mult_general(Ctx& c, const double* K0_16, const double* A1, double c1, const double* A2, double c2,
                       int ncomp, const double* x, double* y) {
    ConstMat m;
    for (int k = 0; k < 16; ++k) m.A[k] = K0_16 ? K0_16[k] : 0.0;
    k_stored_general<<<blocks_for(c.g.nn), BLOCK, 0, c.stream>>>(c.g, m, A1, c1, A2, c2, ncomp, x, y);
    FCFD_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------------------
// Boundary integral F_p = int q (U_d . n) with 2-point Gauss per Dirichlet edge
// (element_calc.py:111-135), gathered per node from a CSR of incident edge slots built at setup
// in bconn order, so each node sums its contributions in segment_sum order.
// ---------------------------------------------------------------------------------------------
__global__ void k_boundary_integral(int nn, const int* __restrict__ off, const int* __restrict__ slot,
                                    const double* __restrict__ geom, const int* __restrict__ ui,
                                    const double* __restrict__ Ud, double* __restrict__ Fp) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const double s = 0.57735026918962576451;  // 1/sqrt(3)
    const double xq[2] = {-s, s};
    double acc = 0.0;
    for (int k = off[n]; k < off[n + 1]; ++k) {
        int ed = slot[k] >> 1, loc = slot[k] & 1;
        double nx = geom[ed * 3 + 0], ny = geom[ed * 3 + 1], h = geom[ed * 3 + 2];
        double u0 = Ud[ui[ed * 4 + 0]], v0 = Ud[ui[ed * 4 + 1]];
        double u1 = Ud[ui[ed * 4 + 2]], v1 = Ud[ui[ed * 4 + 3]];
        double fe = 0.0;
        for (int q = 0; q < 2; ++q) {
            double N0 = (1 - xq[q]) / 2, N1 = (1 + xq[q]) / 2;
            double uq = N0 * u0 + N1 * u1, vq = N0 * v0 + N1 * v1;
            double un = nx * uq + ny * vq;
            double w = h / 2;
            fe += (loc == 0 ? N0 : N1) * w * un;
        }
        acc += fe;
    }
    Fp[n] = acc;
}

void boundary_integral(Ctx& c, const double* U_d, double* F_p) {
    if (c.n_bedge == 0 || U_d == nullptr) {
        launch_fill(c, F_p, 0.0, c.g.nn);
        return;
    }
    k_boundary_integral<<<(c.g.nn + 255) / 256, 256, 0, c.stream>>>(c.g.nn, c.fb_off, c.fb_edge, c.fb_geom,
                                                                       c.fb_ui, U_d, F_p);
    FCFD_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------------------
// mesh: coordinates and connectivity from the same index math the kernels use
// (regular_mesh.py:5-43 / :149-180)
// ---------------------------------------------------------------------------------------------
__global__ void k_mesh_fill(Grid2D g, double dx, double dy, double* coords, int64_t* nconn) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < g.nn && coords) {
        int j = t / g.nnx, i = t - j * g.nnx;
        coords[2 * t] = (double)i * dx;
        coords[2 * t + 1] = (double)j * dy;
    }
    if (t < g.ne && nconn) {
        int cn[4];
        elem_nodes(g, t, cn);
        for (int k = 0; k < 4; ++k) nconn[4 * (size_t)t + k] = cn[k];
    }
}
void mesh_fill(Ctx& c, double* coords, int64_t* nconn) {
    int n = c.g.nn > c.g.ne ? c.g.nn : c.g.ne;
    k_mesh_fill<<<(n + 255) / 256, 256, 0, c.stream>>>(c.g, c.lx / c.g.nex, c.ly / c.g.ney, coords, nconn);
    FCFD_CHECK_LAUNCH();
}

// dense U_d from (gamma_d, U_d)
__global__ void k_scatter_ud(int n, const int64_t* __restrict__ gd, const double* __restrict__ ud, double* full) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) full[gd[k]] = ud[k];
}
void scatter_ud(Ctx& c, const double* U_d) {
    if (c.n_gamma_d == 0) return;
    if (U_d == nullptr) {
        launch_fill(c, c.ud_full, 0.0, c.dim * c.g.nn);
        return;
    }
    k_scatter_ud<<<(c.n_gamma_d + 255) / 256, 256, 0, c.stream>>>(c.n_gamma_d, c.gamma_d_dev, U_d, c.ud_full);
    FCFD_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------------------
// BLAS-1
// ---------------------------------------------------------------------------------------------
struct DotBody {
    const double *x, *y;
    __device__ void operator()(int k, double (&acc)[1]) const { acc[0] += x[k] * y[k]; }
};
struct StoreFin {
    double* out;
    __device__ void operator()(const double (&t)[1]) const { *out = t[0]; }
};

void read_state_ints(Ctx& c, const int* dev, int n, int* host) {
    FCFD_HIP(hipMemcpyAsync(c.h_pinned, dev, n * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    FCFD_HIP(hipStreamSynchronize(c.stream));
    for (int i = 0; i < n; ++i) host[i] = c.h_pinned[i];
}

// *out (device) = x . y: one partial per BLOCK elements into partials[0, blocks_for(n)), then the
// one-block fixed-order finalize.  Enqueued on c.stream, no synchronisation.
void dot_to_device(Ctx& c, const double* x, const double* y, int n, double* partials, double* out) {
    const int nb = blocks_for(n);
    node_kernel<1><<<nb, BLOCK, 0, c.stream>>>(n, nullptr, DotBody{x, y}, partials, nb);
    finalize_kernel<1><<<1, FIN_BLOCK, 0, c.stream>>>(nb, nullptr, StoreFin{out}, partials, nb);
    FCFD_CHECK_LAUNCH();
}

// The generic dot product keeps its own partials buffer, grown to the vector length on demand:
// the solvers' captured graphs hold c.partials, which therefore never moves.
double dot(Ctx& c, const double* x, const double* y, int n) {
    const int nb = blocks_for(n);
    if (nb > c.dot_pstride) {
        FCFD_HIP(hipStreamSynchronize(c.stream));
        for (auto& p : c.allocs)
            if (p == (void*)c.dot_partials) p = nullptr;
        if (c.dot_partials) FCFD_HIP(hipFree(c.dot_partials));
        c.dot_pstride = nb + 1;
        c.dot_partials = dalloc<double>(c, (size_t)c.dot_pstride);
    }
    dot_to_device(c, x, y, n, c.dot_partials, &c.st->dot);
    double out;
    FCFD_HIP(hipMemcpyAsync(c.h_pinned, &c.st->dot, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FCFD_HIP(hipStreamSynchronize(c.stream));
    std::memcpy(&out, c.h_pinned, sizeof(double));
    return out;
}

__global__ void k_axpby(double a, const double* __restrict__ x, double b, double* __restrict__ y, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}
void axpby(Ctx& c, double a, const double* x, double b, double* y, int n) {
    k_axpby<<<(n + 255) / 256, 256, 0, c.stream>>>(a, x, b, y, n);
    FCFD_CHECK_LAUNCH();
}

}  // namespace fcfd
