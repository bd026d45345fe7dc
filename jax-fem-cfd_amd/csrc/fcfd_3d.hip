// 3D (Q1 hex8) four-step timestep on structured box grids.
//
// The same chain as the 2D path (jax_fem_cfd/solvers/standard.py:148-165) with ndim = 3:
//   Ce(U_n, beta = 1/2) per element (element_calc.py:165-185, 8 x 8),
//   BiCGSTAB on (rho/dt) M_c + 1/2 (rho Ce + mu Le) for the 3 velocity components with Dirichlet
//   rows -> identity (standard.py:14-43), U* = U_hat + (dt/rho) G P_n / M, PCG on the pressure
//   Laplacian (standard.py:85-125) preconditioned by jacobi / approx_inv / none or the box-grid
//   V-cycle of fcfd_mg3.hip (the reference has no 3D multigrid, simulation.py:75-76), and
//   U = U* - (dt/rho) G P / M, U[gamma_d] = U_d.
// The reference's consistent_mass_mult fails for ndim = 3 (element_mult.py:26 reads an undefined
// y_result); the mass operator here is the per-component product of its 2D form.
//
// Kernels are owner-computes node gathers as in 2D: a node sums, over its 8 incident elements in
// ascending element order, the row of each element matrix that belongs to it.  Reductions are the
// deterministic two-stage ones with the same finalize steps (device scalar recurrences) as 2D.
#include <algorithm>
#include <cmath>
#include "fcfd_bicgstab.hpp"
#include "fcfd_cg.hpp"

namespace fcfd {

// ------------------------------------------------------------------------------ host set-up
template <class T>
static void upload3(const std::vector<T>& h, T* d) {
    FCFD_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

// hex8 2x2x2 Gauss data of element 0 (shape_functions.py:47-86): N, J = dN/dxi . coords,
// w = det J, dN/dx = J^-1 dN/dxi; then Me, Le, Ge (element_calc.py:9-40).
void hex8_constants(double dx, double dy, double dz, ElemConst3& ec) {
    const double s = 1.0 / std::sqrt(3.0);
    const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                 sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double X[8][3];
    for (int b = 0; b < 8; ++b) {
        X[b][0] = hbx(b) * dx;
        X[b][1] = hby(b) * dy;
        X[b][2] = hbz(b) * dz;
    }
    for (int q = 0; q < 8; ++q) {
        const double xi = sx[q] * s, eta = sy[q] * s, zeta = sz[q] * s;
        double dxi[3][8];
        for (int I = 0; I < 8; ++I) {
            const double fx = 1 + xi * sx[I], fy = 1 + eta * sy[I], fz = 1 + zeta * sz[I];
            ec.N[q * 8 + I] = fx * fy * fz / 8;
            dxi[0][I] = sx[I] * fy * fz / 8;
            dxi[1][I] = fx * sy[I] * fz / 8;
            dxi[2][I] = fx * fy * sz[I] / 8;
        }
        double J[3][3];
        for (int r = 0; r < 3; ++r)
            for (int col = 0; col < 3; ++col) {
                double a = 0.0;
                for (int I = 0; I < 8; ++I) a += dxi[r][I] * X[I][col];
                J[r][col] = a;
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
        ec.wq[q] = det;
        for (int d = 0; d < 3; ++d)
            for (int I = 0; I < 8; ++I)
                ec.dN[d][q * 8 + I] = Ji[d][0] * dxi[0][I] + Ji[d][1] * dxi[1][I] + Ji[d][2] * dxi[2][I];
    }
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            double m = 0.0, l = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
            for (int q = 0; q < 8; ++q) {
                const double w = ec.wq[q], Na = ec.N[q * 8 + a], Nb = ec.N[q * 8 + b];
                m += w * Na * Nb;
                l += w * (ec.dN[0][q * 8 + a] * ec.dN[0][q * 8 + b] + ec.dN[1][q * 8 + a] * ec.dN[1][q * 8 + b] +
                          ec.dN[2][q * 8 + a] * ec.dN[2][q * 8 + b]);
                g0 += w * Na * ec.dN[0][q * 8 + b];
                g1 += w * Na * ec.dN[1][q * 8 + b];
                g2 += w * Na * ec.dN[2][q * 8 + b];
            }
            ec.Me[a * 8 + b] = m;
            ec.Le[a * 8 + b] = l;
            ec.G[0][a * 8 + b] = g0;
            ec.G[1][a * 8 + b] = g1;
            ec.G[2][a * 8 + b] = g2;
            ec.Leoff[a * 8 + b] = a == b ? 0.0 : l;
        }
}

// out[n] = sum over the incident elements of per_corner[the node's corner], in element order
void node_sums3(const Grid3D& g, const double (&per_corner)[8], std::vector<double>& out) {
    out.assign(g.nn, 0.0);
    for (int ke = 0; ke < g.nez; ++ke)
        for (int je = 0; je < g.ney; ++je)
            for (int ie = 0; ie < g.nex; ++ie) {
                const int n0 = ke * g.nxy + je * g.nnx + ie;
                for (int b = 0; b < 8; ++b) out[n0 + hbz(b) * g.nxy + hby(b) * g.nnx + hbx(b)] += per_corner[b];
            }
}

// quad4 2x2 Gauss weights of a face from the (z, y) coordinates of its nodes, as
// u_dot_n_boundary_integral_2d evaluates every face (element_calc.py:137-163: the face
// coordinates are always columns [2, 1], so faces of constant y or z have det J = 0 and add
// nothing).  Returns W[a][b] = sum_q w_q N_qa N_qb.
static void face_weights(const double (&P)[4][2], double (&W)[16]) {
    const double s = 1.0 / std::sqrt(3.0);
    const double xi[4] = {-s, s, s, -s}, eta[4] = {-s, -s, s, s};
    for (int k = 0; k < 16; ++k) W[k] = 0.0;
    for (int q = 0; q < 4; ++q) {
        const double N[4] = {(1 - xi[q]) * (1 - eta[q]) / 4, (1 + xi[q]) * (1 - eta[q]) / 4,
                             (1 + xi[q]) * (1 + eta[q]) / 4, (1 - xi[q]) * (1 + eta[q]) / 4};
        const double dxi[4] = {-(1 - eta[q]) / 4, (1 - eta[q]) / 4, (1 + eta[q]) / 4, -(1 + eta[q]) / 4};
        const double deta[4] = {-(1 - xi[q]) / 4, -(1 + xi[q]) / 4, (1 + xi[q]) / 4, (1 - xi[q]) / 4};
        double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
        for (int I = 0; I < 4; ++I) {
            J00 += dxi[I] * P[I][0]; J01 += dxi[I] * P[I][1];
            J10 += deta[I] * P[I][0]; J11 += deta[I] * P[I][1];
        }
        const double w = J00 * J11 - J01 * J10;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) W[a * 4 + b] += w * N[a] * N[b];
    }
}

void build_ctx3(Ctx& c, const fcfd_mesh_desc* m, const fcfd_solver_desc* s) {
    FCFD_REQUIRE(m->nex >= 1 && m->ney >= 1 && m->nez >= 1, "bad element counts");
    FCFD_REQUIRE(!m->periodic, "3D periodic meshes are not supported (the reference rejects them)");
    FCFD_REQUIRE(c.precond == FCFD_PC_NONE || c.precond == FCFD_PC_JACOBI || c.precond == FCFD_PC_APPROX_INV ||
                     c.precond == FCFD_PC_MULTIGRID,
                 "3D pressure preconditioner must be none, jacobi, approx_inv or multigrid");
    FCFD_REQUIRE(c.ce_format == FCFD_CE_F64 || c.ce_format == FCFD_CE_F32, "3D Ce formats: FP64 or FP32");
    FCFD_REQUIRE((int64_t)(m->nex + 1) * (m->ney + 1) * (m->nez + 1) * 3 < (int64_t)1 << 31, "grid too large");
    c.dim = 3;
    c.g3 = make_grid3(m->nex, m->ney, m->nez);
    const Grid3D& g = c.g3;
    c.g = Grid2D{0, 0, 0, 0, g.nn, g.ne, 0};
    c.lx = m->lx; c.ly = m->ly; c.lz = m->lz;
    const int nn = g.nn;
    hex8_constants(c.lx / g.nex, c.ly / g.ney, c.lz / g.nez, c.ec3);
    c.ec3_dev = dalloc<ElemConst3>(c, 1);
    FCFD_HIP(hipMemcpy(c.ec3_dev, &c.ec3, sizeof(ElemConst3), hipMemcpyHostToDevice));

    // lumped mass (element_calc.py:9-25), tiled per component
    {
        double rs[8];
        for (int a = 0; a < 8; ++a) {
            rs[a] = 0.0;
            for (int b = 0; b < 8; ++b) rs[a] += c.ec3.Me[a * 8 + b];
        }
        std::vector<double> M;
        node_sums3(g, rs, M);
        M.resize(3 * (size_t)nn);
        std::copy(M.begin(), M.begin() + nn, M.begin() + nn);
        std::copy(M.begin(), M.begin() + nn, M.begin() + 2 * (size_t)nn);
        c.Mlump = dalloc<double>(c, 3 * (size_t)nn);
        upload3(M, c.Mlump);
    }
    // one level: L_diag (element_calc.py:99-109) and the preconditioner scratch
    c.nlev = 1;
    c.lv.resize(1);
    {
        Level& L = c.lv[0];
        L.g = c.g;
        double dg[8];
        for (int a = 0; a < 8; ++a) dg[a] = c.ec3.Le[a * 8 + a];
        std::vector<double> Ld;
        node_sums3(g, dg, Ld);
        L.Ld = dalloc<double>(c, nn);
        upload3(Ld, L.Ld);
        L.x = dalloc<double>(c, nn);
        L.y = dalloc<double>(c, nn);
    }
    if (c.precond == FCFD_PC_MULTIGRID) build_mg3(c, s);
    // pressure pins (outlets, else node 0)
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
        upload3(pin, c.ppin);
    }
    // Dirichlet DOFs (3 components)
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
        upload3(mask, c.dmask);
        c.gamma_d_dev = dalloc<int64_t>(c, gd.size());
        upload3(gd, c.gamma_d_dev);
        c.ud_full = dalloc<double>(c, 3 * (size_t)nn);
        FCFD_HIP(hipMemset(c.ud_full, 0, sizeof(double) * 3 * nn));
    }
    // Dirichlet faces -> per-node gather lists in bconn order (segment_sum order)
    if (s->bconn && s->nvec && s->n_bedge > 0 && !gd.empty()) {
        const int nf = (int)s->n_bedge;
        c.n_bedge = nf;
        std::vector<std::vector<int>> lists(nn);
        std::vector<double> W(16 * (size_t)nf), geom(3 * (size_t)nf);
        std::vector<int> ui(12 * (size_t)nf);
        const double dx = c.lx / g.nex, dy = c.ly / g.ney, dz = c.lz / g.nez;
        auto sorted_index = [&](int64_t v) { return (int)(std::lower_bound(gd.begin(), gd.end(), v) - gd.begin()); };
        for (int f = 0; f < nf; ++f) {
            double P[4][2];
            for (int a = 0; a < 4; ++a) {
                const int64_t v = s->bconn[4 * (size_t)f + a];
                FCFD_REQUIRE(v >= 0 && v < nn, "bconn out of range");
                const int kk = (int)(v / g.nxy), r = (int)(v % g.nxy);
                P[a][0] = kk * dz;            // z
                P[a][1] = (r / g.nnx) * dy;   // y
                for (int d = 0; d < 3; ++d) {
                    const int idx = sorted_index(v + (int64_t)d * nn);
                    FCFD_REQUIRE(idx < (int)gd.size() && gd[idx] == v + (int64_t)d * nn,
                                 "boundary node is not a Dirichlet DOF");
                    ui[12 * (size_t)f + 3 * a + d] = idx;
                }
                lists[v].push_back(f * 4 + a);
            }
            (void)dx;
            double w[16];
            face_weights(P, w);
            for (int k = 0; k < 16; ++k) W[16 * (size_t)f + k] = w[k];
            for (int d = 0; d < 3; ++d) geom[3 * (size_t)f + d] = (double)s->nvec[3 * (size_t)f + d];
        }
        std::vector<int> off(nn + 1, 0), slots;
        for (int n = 0; n < nn; ++n) {
            off[n + 1] = off[n] + (int)lists[n].size();
            slots.insert(slots.end(), lists[n].begin(), lists[n].end());
        }
        c.fb_off = dalloc<int>(c, nn + 1);
        upload3(off, c.fb_off);
        c.fb_edge = dalloc<int>(c, slots.size());
        upload3(slots, c.fb_edge);
        c.fb_w = dalloc<double>(c, W.size());
        upload3(W, c.fb_w);
        c.fb_geom = dalloc<double>(c, geom.size());
        upload3(geom, c.fb_geom);
        c.fb_ui = dalloc<int>(c, ui.size());
        upload3(ui, c.fb_ui);
    }
    // element matrices and solver work (3 components)
    const size_t ne = g.ne;
    if (c.ce_format == FCFD_CE_F64) c.ce = dalloc<double>(c, 64 * ne);
    else c.ce = dalloc<float>(c, 64 * ne);
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
    const dim3 tiles = s3t_tiles(c.g3);  // blocks of the box-tiled stencil kernels
    c.pstride = std::max(blocks_for((int)n3), (int)(tiles.x * tiles.y * tiles.z)) + 1;
    c.partials = dalloc<double>(c, (size_t)MAX_RED * c.pstride);
    c.st = dalloc<DevState>(c, 1);
    FCFD_HIP(hipMemset(c.st, 0, sizeof(DevState)));
    FCFD_HIP(hipHostMalloc((void**)&c.h_pinned, 64, hipHostMallocDefault));
    FCFD_HIP(hipDeviceSynchronize());
}

// ------------------------------------------------------------------------------ Ce build
// Ce[a][b] = sum_q w N_qa (U_q . grad N_qb + beta div U_q N_qb)   (element_calc.py:165-185)
template <class S>
__global__ __launch_bounds__(BLOCK) void k_build_ce3(Grid3D g, const ElemConst3* __restrict__ ec,
                                                    const double* __restrict__ U, double beta, S store) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (e >= g.ne) return;
    int cn[8];
    elem_nodes3(g, e, cn);
    double ul[3][8];
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int d = 0; d < 3; ++d) ul[d][b] = U[cn[b] + (size_t)d * g.nn];
    double A[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) A[k] = 0.0;
    for (int q = 0; q < 8; ++q) {
        double uq[3] = {0.0, 0.0, 0.0}, div = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const double Nb = ec->N[q * 8 + b];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                uq[d] += Nb * ul[d][b];
                div += ec->dN[d][q * 8 + b] * ul[d][b];
            }
        }
        const double w = ec->wq[q];
        double adv[8];
#pragma unroll
        for (int b = 0; b < 8; ++b)
            adv[b] = uq[0] * ec->dN[0][q * 8 + b] + uq[1] * ec->dN[1][q * 8 + b] + uq[2] * ec->dN[2][q * 8 + b] +
                     beta * div * ec->N[q * 8 + b];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const double wNa = w * ec->N[q * 8 + a];
#pragma unroll
            for (int b = 0; b < 8; ++b) A[a * 8 + b] += wNa * adv[b];
        }
    }
#pragma unroll
    for (int k = 0; k < 64; ++k) store.st(k, e, A[k]);
}

// element API: Ce (ne, 8, 8) in the reference's row-major layout, FP64
struct Ce3Api {
    double* c;
    __device__ __forceinline__ void st(int k, int e, double v) const { c[(size_t)e * 64 + k] = v; }
};
void build_ce3_api(Ctx& c, const double* U, double beta, double* Ce) {
    k_build_ce3<<<blocks_for(c.g3.ne), BLOCK, 0, c.stream>>>(c.g3, c.ec3_dev, U, beta, Ce3Api{Ce});
    FCFD_CHECK_LAUNCH();
}

static void build_ce3(Ctx& c, const double* U, double beta) {
    const int nb = blocks_for(c.g3.ne);
    if (c.ce_format == FCFD_CE_F64)
        k_build_ce3<<<nb, BLOCK, 0, c.stream>>>(c.g3, c.ec3_dev, U, beta, Ce3F64{(double*)c.ce, c.g3.ne});
    else
        k_build_ce3<<<nb, BLOCK, 0, c.stream>>>(c.g3, c.ec3_dev, U, beta, Ce3F32{(float*)c.ce, c.g3.ne});
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ momentum
// (cM Me + cL Le + cC Ce_e), one row of element e
template <class S>
struct MomCoef3 {
    Grid3D g;
    const ElemConst3* __restrict__ ec;
    S ce;
    double cM, cL, cC;
    __device__ __forceinline__ void row(int a, int e, double (&cf)[8]) const {
        double v[8];
        ce.ld_row(a, e, v);
#pragma unroll
        for (int b = 0; b < 8; ++b) cf[b] = cM * ec->Me[a * 8 + b] + cL * ec->Le[a * 8 + b] + cC * v[b];
    }
};

template <int NC, class Op>
__device__ __forceinline__ void op_apply3(const Op& op, const Nbr3& s, const double (&X)[NC][27], double (&y)[NC]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) y[c] = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double cf[8];
        op.row(k3Row(k), s.e[k], cf);
        double t[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) t[c] = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c) t[c] += cf[b] * X[c][k3Slot(k, b)];
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] += s.has[k] ? t[c] : 0.0;
    }
}

// rhs = R U_n - G P_n + F (Dirichlet rows U_d), x0 = U_n with Dirichlet values, r0 = rhs - A x0,
// rhat = p = q = r0; reductions b.b, r0.r0 (the 2D BInit with 3 components)
template <class Op>
struct BInit3 {
    Op lhs, rhs;
    const double *xn, *ud, *Pn, *F;
    const uint8_t* mask;
    double *x, *r, *rhat, *p, *q;
    __device__ void operator()(int n, double (&acc)[2]) const {
        const int nn = lhs.g.nn;
        Nbr3 s;
        node_nbr3(lhs.g, n, s);
        double X[3][27], X0[3][27];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double* xc = xn + c * nn;
            load27(s, [&](int k) { return xc[k]; }, X[c]);
            if (mask) {
                const double* uc = ud + c * nn;
                const uint8_t* mc = mask + c * nn;
                load27(s, [&](int k) { const double a = uc[k], b = xc[k]; return mc[k] ? a : b; }, X0[c]);
            } else {
#pragma unroll
                for (int t = 0; t < 27; ++t) X0[c][t] = X[c][t];
            }
        }
        double b[3], ax[3];
        op_apply3<3>(rhs, s, X, b);
        {
            double XP[27];
            load27(s, [&](int k) { return Pn[k]; }, XP);
            const ElemConst3* ec = lhs.ec;
            b[0] -= ebe3(s, ec->G[0], XP);
            b[1] -= ebe3(s, ec->G[1], XP);
            b[2] -= ebe3(s, ec->G[2], XP);
            if (F) { b[0] += F[n]; b[1] += F[n + nn]; b[2] += F[n + 2 * nn]; }
        }
        op_apply3<3>(lhs, s, X0, ax);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = n + c * nn;
            const bool m = mask && mask[i];
            const double bc = m ? ud[i] : b[c];
            const double a = m ? X0[c][S3C] : ax[c];
            const double rc = bc - a;
            x[i] = X0[c][S3C];
            r[i] = rc; rhat[i] = rc; p[i] = rc; q[i] = rc;
            acc[0] += bc * bc;
            acc[1] += rc * rc;
        }
    }
};

// LDS-staged forms: the on-the-fly vector (p or s, 3 components) staged per block; the element
// rows are applied to the staged stencil directly, one element at a time (a rolled loop keeps the
// live set to one row; the element id and validity are recomputed per element so that no private
// array is indexed at run time)
template <class Op, class SV>
__device__ __forceinline__ void op_apply3s(const Op& op, const SV& sv, int n, double (&y)[3]) {
    const Grid3D& g = op.g;
    const int kz = n / g.nxy, rr = n - kz * g.nxy;
    const int jy = rr / g.nnx, ix = rr - jy * g.nnx;
#pragma unroll
    for (int c = 0; c < 3; ++c) y[c] = 0.0;
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
        const int bx = k & 1, by = (k >> 1) & 1, bz = k >> 2;
        const int ie = ix - 1 + bx, je = jy - 1 + by, ke = kz - 1 + bz;
        const bool has = ie >= 0 && ie < g.nex && je >= 0 && je < g.ney && ke >= 0 && ke < g.nez;
        const int e = has ? (ke * g.ney + je) * g.nex + ie : 0;
        const int a = hloc(1 - bx, 1 - by, 1 - bz);
        double cf[8];
        op.row(a, e, cf);
        double t[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int slot = ((hbz(b) + bz) * 3 + (hby(b) + by)) * 3 + (hbx(b) + bx);
#pragma unroll
            for (int c = 0; c < 3; ++c) t[c] += cf[b] * sv.at(c, n, slot);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) y[c] += has ? t[c] : 0.0;
    }
}

template <class Op>
struct BK1S3 {
    static constexpr int NV = 3;
    Op op;
    const uint8_t* mask;
    const double *r, *pold, *qold, *rhat;
    double *pnew, *qnew;
    const DevState* st;
    __device__ __forceinline__ void stage(int id, double (&v)[3]) const {
        const double beta = st->b_beta, omega = st->b_omega;
        const int nn = op.g.nn;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = id + c * nn;
            v[c] = r[i] + beta * (pold[i] - omega * qold[i]);
        }
    }
    template <class SV>
    __device__ void operator()(int n, const SV& sv, double (&acc)[1]) const {
        const int nn = op.g.nn;
        double y[3];
        op_apply3s(op, sv, n, y);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = n + c * nn;
            const double pc = sv.at(c, n, S3C);
            const double yc = (mask && mask[i]) ? pc : y[c];
            pnew[i] = pc;
            qnew[i] = yc;
            acc[0] += rhat[i] * yc;
        }
    }
};

template <class Op>
struct BK2S3 {
    static constexpr int NV = 3;
    Op op;
    const uint8_t* mask;
    const double *r, *q;
    double *sv_, *tv;
    const DevState* st;
    __device__ __forceinline__ void stage(int id, double (&v)[3]) const {
        const double alpha = st->b_alpha;
        const int nn = op.g.nn;
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = r[id + c * nn] - alpha * q[id + c * nn];
    }
    template <class SV>
    __device__ void operator()(int n, const SV& sv, double (&acc)[3]) const {
        const int nn = op.g.nn;
        double y[3];
        op_apply3s(op, sv, n, y);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = n + c * nn;
            const double sc = sv.at(c, n, S3C);
            const double yc = (mask && mask[i]) ? sc : y[c];
            sv_[i] = sc;
            tv[i] = yc;
            acc[0] += sc * sc;
            acc[1] += yc * sc;
            acc[2] += yc * yc;
        }
    }
};

template <int NR, class Body, class Fin>
static void launch_stencil3(Ctx& c, const Grid3D& g, const int* active, const Body& body, const Fin& fin) {
    const dim3 t = s3t_tiles(g);
    const int nb = (int)(t.x * t.y * t.z);
    if (NR > 0 && nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
    stencil3t_kernel<NR><<<nb, BLOCK, 0, c.stream>>>(g, (int)t.x, (int)t.y, active, body, c.partials, c.pstride);
    if constexpr (NR > 0)
        finalize_kernel<NR><<<1, FIN_BLOCK, 0, c.stream>>>(nb, active, fin, c.partials, c.pstride);
}

template <class S>
static int solve_momentum3_t(Ctx& c, S ce, const double* U_n, const double* P_n, double rho, double mu, double dt,
                             const double* F, double* U_hat) {
    const double theta = 0.5;
    const Grid3D& g = c.g3;
    const int nn = g.nn;
    MomCoef3<S> lhs{g, c.ec3_dev, ce, rho / dt, theta * mu, theta * rho};
    MomCoef3<S> rhs{g, c.ec3_dev, ce, rho / dt, -(1.0 - theta) * mu, -(1.0 - theta) * rho};
    const int maxit = c.maxit_mom > 0 ? c.maxit_mom : 10 * 3 * nn;
    launch_node<2>(c, nn, nullptr,
                   BInit3<MomCoef3<S>>{lhs, rhs, U_n, c.ud_full, P_n, F, c.dmask, c.m_x, c.m_r, c.m_rhat, c.m_p[0],
                                       c.m_q[0]},
                   BInitFin{c.st, c.tol, maxit});
    FCFD_CHECK_LAUNCH();
    const int* act = &c.st->b_active;
    auto pair = [&]() {
        for (int o = 0; o < 2; ++o) {
            const int nw = o ^ 1;
            launch_stencil3<1>(c, g, act,
                               BK1S3<MomCoef3<S>>{lhs, c.dmask, c.m_r, c.m_p[o], c.m_q[o], c.m_rhat, c.m_p[nw],
                                                  c.m_q[nw], c.st},
                               BK1Fin{c.st});
            launch_stencil3<3>(c, g, act, BK2S3<MomCoef3<S>>{lhs, c.dmask, c.m_r, c.m_q[nw], c.m_s, c.m_t, c.st},
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

int solve_momentum3(Ctx& c, const double* U_n, const double* P_n, double rho, double mu, double dt, const double* F,
                    double* U_hat) {
    if (c.ce_format == FCFD_CE_F64)
        return solve_momentum3_t(c, Ce3F64{(double*)c.ce, c.g3.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
    return solve_momentum3_t(c, Ce3F32{(float*)c.ce, c.g3.ne}, U_n, P_n, rho, mu, dt, F, U_hat);
}

int solve_momentum3_from(Ctx& c, const double* U_n, const double* P_n, double rho, double mu, double dt,
                         const double* F, double* U_hat) {
    build_ce3(c, U_n, 0.5);
    return solve_momentum3(c, U_n, P_n, rho, mu, dt, F, U_hat);
}

// ------------------------------------------------------------------------------ projection
// U* = U_hat + (dt/rho) G P_n / M   (standard.py:156); sgn = -1: U = U* - (dt/rho) G P / M with
// U[gamma_d] = U_d   (standard.py:160-163)
struct Proj3S {
    static constexpr int NV = 1;
    Grid3D g;
    const ElemConst3* ec;
    const double *Uin, *P, *M, *ud;
    const uint8_t* mask;
    double* Uout;
    double c;
    __device__ __forceinline__ void stage(int id, double (&v)[1]) const { v[0] = P[id]; }
    template <class SV>
    __device__ void operator()(int n, const SV& sv, double (&)[1]) const {
        Nbr3 s;
        node_elems3(g, n, s);
        double X[27];
        view27(sv, 0, n, X);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int i = n + d * g.nn;
            const double gd = ebe3(s, ec->G[d], X);
            double u = Uin[i] + (c * gd) / M[i];
            if (mask && mask[i]) u = ud[i];
            Uout[i] = u;
        }
    }
};

// ------------------------------------------------------------------------------ pressure
// rhs = rho/dt (D U* - F_p), pinned rows zeroed; sum for the mean projection
struct Prhs3Body {
    Grid3D g; const ElemConst3* ec; const double* Us; const double* Fp; const uint8_t* pin; double* rhs; double c;
    __device__ void operator()(int n, double (&acc)[1]) const {
        Nbr3 s;
        node_nbr3(g, n, s);
        double div = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double X[27];
            const double* u = Us + (size_t)d * g.nn;
            load27(s, [&](int k) { return u[k]; }, X);
            div += ebe3<true>(s, ec->G[d], X);
        }
        double v = c * (div - (Fp ? Fp[n] : 0.0));
        if (pin && pin[n]) v = 0.0;
        rhs[n] = v;
        acc[0] += v;
    }
};

struct CgInit3 {
    Grid3D g; const double* Le; const uint8_t* pin; const double* b; const double* bmean; const double* x0in;
    int zero_pins; double* x; double* r;
    __device__ void operator()(int n, double (&acc)[3]) const {
        Nbr3 s;
        node_nbr3(g, n, s);
        double X[27];
        if (zero_pins && pin)
            load27(s, [&](int k) { const double v = x0in[k]; return pin[k] ? 0.0 : v; }, X);
        else
            load27(s, [&](int k) { return x0in[k]; }, X);
        const bool pn = pin && pin[n];
        const double ax = pn ? 0.0 : ebe3(s, Le, X);
        const double bn = b[n] - (bmean ? *bmean : 0.0);
        const double rn = bn - ax;
        x[n] = X[S3C];
        r[n] = rn;
        acc[0] += bn * bn;
        acc[1] += rn * rn;
        acc[2] += rn;
    }
};

// two-term Neumann-series approximate inverse (standard.py:225-234)
struct ApproxInv3a {
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; double* d1l;
    __device__ void operator()(int n, double (&)[1]) const {
        Nbr3 s;
        node_nbr3(g, n, s);
        double X[27];
        load27(s, [&](int k) { return r[k] / Ld[k]; }, X);
        const double v = ebe3(s, Le, X) / Ld[n];
        d1l[n] = (pin && pin[n]) ? 0.0 : v;
    }
};
struct ApproxInv3b {
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; const double* d1l; double* z;
    __device__ void operator()(int n, double (&)[1]) const {
        Nbr3 s;
        node_nbr3(g, n, s);
        double X[27];
        load27(s, [&](int k) { return d1l[k]; }, X);
        double l2 = ebe3(s, Le, X);
        if (pin && pin[n]) l2 = 0.0;
        const double d1x = r[n] / Ld[n];
        z[n] = (3.0 * d1x - 3.0 * X[S3C]) + l2 / Ld[n];
    }
};

// LDS-staged forms of CgK1_3 / ApproxInv3a / ApproxInv3b (the latter fused with the CG update
// z_new = M r, partial r.z_new that follows it)
struct CgK1S3 {
    static constexpr int NV = 1;
    Grid3D g; const double* Le; const uint8_t* pin; const double* z; const double* pold; double* pnew; double* Ap;
    const DevState* st;
    __device__ __forceinline__ void stage(int id, double (&v)[1]) const { v[0] = z[id] + st->c_beta * pold[id]; }
    template <class SV>
    __device__ void operator()(int n, const SV& sv, double (&acc)[1]) const {
        Nbr3 s;
        node_elems3(g, n, s);
        double X[27];
        view27(sv, 0, n, X);
        const double a = (pin && pin[n]) ? 0.0 : ebe3(s, Le, X);
        pnew[n] = X[S3C];
        Ap[n] = a;
        acc[0] += X[S3C] * a;
    }
};
// ApproxInv3a, LDS-staged, with CgK2's x / r update (jax_iterative_solvers.py:108-109) fused in:
// the window stages r_new / L_diag with r_new = r - alpha Ap formed as CgK2 forms it, and the block stores
// x + alpha p and r_new of its own nodes (into rout: rin stays intact, the neighbouring blocks
// stage it) with the partial of r_new.r_new
struct ApproxInv3aK2S {
    static constexpr int NV = 1;
    Grid3D g; const double* Le; const double* rin; const double* Ap; const double* p; double* x; double* rout;
    const double* Ld; const uint8_t* pin; double* d1l; const DevState* st;
    __device__ __forceinline__ void stage(int id, double (&v)[1]) const {
        const double rk = rin[id] - st->c_alpha * Ap[id];
        v[0] = rk / Ld[id];
    }
    template <class SV>
    __device__ void operator()(int n, const SV& sv, double (&acc)[1]) const {
        // the stencil before any store: a store ahead of it would force Le's uniform loads off the
        // scalar path (the compiler cannot rule out that x or rout alias Le)
        Nbr3 s;
        node_elems3(g, n, s);
        double X[27];
        view27(sv, 0, n, X);
        const double v = ebe3(s, Le, X) / Ld[n];
        const double a = st->c_alpha;
        const double xn = x[n] + a * p[n];
        const double rk = rin[n] - a * Ap[n];
        x[n] = xn;
        rout[n] = rk;
        acc[0] += rk * rk;
        d1l[n] = (pin && pin[n]) ? 0.0 : v;
    }
};
struct ApproxInv3bK3S {
    static constexpr int NV = 1;
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; const double* d1l;
    double* z;
    __device__ __forceinline__ void stage(int id, double (&v)[1]) const { v[0] = d1l[id]; }
    template <class SV>
    __device__ void operator()(int n, const SV& sv, double (&acc)[2]) const {
        Nbr3 s;
        node_elems3(g, n, s);
        double X[27];
        view27(sv, 0, n, X);
        double l2 = ebe3(s, Le, X);
        if (pin && pin[n]) l2 = 0.0;
        const double d1x = r[n] / Ld[n];
        const double zk = (3.0 * d1x - 3.0 * X[S3C]) + l2 / Ld[n];
        z[n] = zk;
        acc[0] += r[n] * zk;
    }
};


static const double* precondition3(Ctx& c, const int* active) {
    const int nn = c.g3.nn;
    switch (c.precond) {
        case FCFD_PC_NONE: return c.c_r;
        case FCFD_PC_JACOBI:
            launch_node<0>(c, nn, active, JacobiPrec{c.c_r, c.lv[0].Ld, c.lv[0].y}, NoFin{});
            return c.lv[0].y;
        case FCFD_PC_MULTIGRID: return vcycle3(c, c.c_r, active);
        case FCFD_PC_APPROX_INV:
            launch_node<0>(c, nn, active, ApproxInv3a{c.g3, c.ec3_dev->Le, c.c_r, c.lv[0].Ld, c.ppin, c.lv[0].x},
                           NoFin{});
            launch_node<0>(c, nn, active,
                           ApproxInv3b{c.g3, c.ec3_dev->Le, c.c_r, c.lv[0].Ld, c.ppin, c.lv[0].x, c.lv[0].y}, NoFin{});
            return c.lv[0].y;
    }
    throw Error(FCFD_ERR_UNSUPPORTED, "3D pressure preconditioner not supported");
}

// PCG (jax_iterative_solvers.py:87-120); solution in c.c_x
static int pcg3(Ctx& c, const double* b, const double* bmean, const double* x0in, int zero_pins, double tol,
                int maxit, bool* converged) {
    const Grid3D& g = c.g3;
    const int nn = g.nn;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    const double* Le = c.ec3_dev->Le;
    launch_node<3>(c, nn, nullptr, CgInit3{g, Le, pin, b, bmean, x0in, zero_pins, c.c_x, c.c_r},
                   CgInitFin{c.st, tol, maxit, nn, 0});
    const int* act = &c.st->c_active;
    const int neu = c.precond == FCFD_PC_MULTIGRID && !c.set_zeroP;
    const double* y0 = precondition3(c, act);
    launch_node<1>(c, nn, act, CgGamma0<double>{c.c_r, y0, c.c_z[0], c.c_p[0], c.st, neu}, CgGamma0Fin{c.st});
    FCFD_CHECK_LAUNCH();
    const double* Ld = c.lv[0].Ld;
    auto pair = [&]() {
        for (int o = 0; o < 2; ++o) {
            const int nw = o ^ 1;
            launch_stencil3<1>(c, g, act, CgK1S3{g, Le, pin, c.c_z[o], c.c_p[o], c.c_p[nw], c.c_Ap, c.st},
                               CgK1Fin{c.st});
            // CgK2's partials of r.r and sum r go to slots 2, 3: without the pure-Neumann V-cycle
            // (which reads the mean of r) nothing reads r.r before the z-step, whose finalize
            // takes it (CgK3RFin: CgK2Fin's c_rr, then CgK3Fin; the same sums in the same order)
            const int nb2 = blocks_for(nn);
            auto k2 = [&]() {
                node_kernel<2><<<nb2, BLOCK, 0, c.stream>>>(nn, act, CgK2{c.c_p[nw], c.c_Ap, c.c_x, c.c_r, c.st},
                                                             c.partials + 2 * (size_t)c.pstride, c.pstride);
            };
            if (c.precond == FCFD_PC_MULTIGRID && neu) {
                launch_node<2>(c, nn, act, CgK2{c.c_p[nw], c.c_Ap, c.c_x, c.c_r, c.st}, CgK2Fin{c.st, nn});
                const double* y = precondition3(c, act);
                const int flex = c.coarse_inv32 != nullptr;
                launch_node<2>(c, nn, act, CgK3<double>{c.c_r, y, c.c_z[nw], c.c_z[o], c.st, neu, flex},
                               CgK3Fin{c.st, flex});
            } else if (c.precond == FCFD_PC_MULTIGRID) {
                k2();
                const double* y = precondition3(c, act);
                const int flex = c.coarse_inv32 != nullptr;
                node_kernel<2><<<nb2, BLOCK, 0, c.stream>>>(
                    nn, act, CgK3<double>{c.c_r, y, c.c_z[nw], c.c_z[o], c.st, 0, flex}, c.partials, c.pstride);
                finalize2_kernel<2, 1><<<1, FIN_BLOCK, 0, c.stream>>>(nb2, nb2, act, CgK3RFin{c.st, flex},
                                                                      c.partials, c.pstride);
            } else if (c.precond == FCFD_PC_APPROX_INV) {
                // the x / r update rides on the first stage (its r.r partials in slot 2); r
                // alternates between c_r (parity 0) and c_r2
                const double* rin = o == 0 ? c.c_r : c.c_r2;
                double* rout = o == 0 ? c.c_r2 : c.c_r;
                const dim3 t = s3t_tiles(g);
                const int nb3 = (int)(t.x * t.y * t.z);
                if (nb3 > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
                stencil3t_kernel<1><<<nb3, BLOCK, 0, c.stream>>>(
                    g, (int)t.x, (int)t.y, act,
                    ApproxInv3aK2S{g, Le, rin, c.c_Ap, c.c_p[nw], c.c_x, rout, Ld, c.ppin, c.lv[0].x, c.st},
                    c.partials + 2 * (size_t)c.pstride, c.pstride);
                stencil3t_kernel<2><<<nb3, BLOCK, 0, c.stream>>>(
                    g, (int)t.x, (int)t.y, act, ApproxInv3bK3S{g, Le, rout, Ld, c.ppin, c.lv[0].x, c.c_z[nw]},
                    c.partials, c.pstride);
                finalize2_kernel<2, 1><<<1, FIN_BLOCK, 0, c.stream>>>(nb3, nb3, act, CgK3RFin{c.st, 0}, c.partials,
                                                                      c.pstride);
            } else {
                launch_node<2>(c, nn, act,
                               CgK23{c.c_p[nw], c.c_Ap, c.precond == FCFD_PC_JACOBI ? Ld : nullptr, c.c_x, c.c_r,
                                     c.c_z[nw], c.st},
                               CgK23Fin{c.st});
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

int pressure_cg3(Ctx& c, const double* b, double* x, double tol, int maxit, bool* converged) {
    const int it = pcg3(c, b, nullptr, x, 0, tol, maxit, converged);
    FCFD_HIP(hipMemcpyAsync(x, c.c_x, sizeof(double) * c.g3.nn, hipMemcpyDeviceToDevice, c.stream));
    return it;
}

int solve_pressure3(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                    double* P_out) {
    const int nn = c.g3.nn;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    launch_node<1>(c, nn, nullptr, Prhs3Body{c.g3, c.ec3_dev, U_star, F_p, pin, c.c_rhs, rho / dt},
                   MeanFin{&c.st->mean_rhs, nullptr, nn});
    const int maxit = c.maxit_p > 0 ? c.maxit_p : 1000;
    bool conv;
    const int it = pcg3(c, c.c_rhs, c.set_zeroP ? nullptr : &c.st->mean_rhs, P_n, c.set_zeroP, c.tol, maxit, &conv);
    if (!c.set_zeroP) launch_node<1>(c, nn, nullptr, SumXBody{c.c_x}, MeanFin{&c.st->mean_p, nullptr, nn});
    launch_node<0>(c, nn, nullptr, PoutBody{c.c_x, pin, P_out, c.st, !c.set_zeroP}, NoFin{});
    FCFD_CHECK_LAUNCH();
    return it;
}

// ------------------------------------------------------------------------------ boundary integral
// F_p[n] = sum over the node's Dirichlet-face slots (bconn order) of
//          sum_b W_f[a][b] (n_f . U_d(b))      (element_calc.py:137-163)
__global__ void k_boundary_integral3(int nn, const int* __restrict__ off, const int* __restrict__ slot,
                                     const double* __restrict__ W, const double* __restrict__ nv,
                                     const int* __restrict__ ui, const double* __restrict__ U_d, double* F_p) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    double acc = 0.0;
    for (int k = off[n]; k < off[n + 1]; ++k) {
        const int f = slot[k] >> 2, a = slot[k] & 3;
        double t = 0.0;
        for (int b = 0; b < 4; ++b) {
            const int* u = ui + 12 * (size_t)f + 3 * b;
            const double un = nv[3 * f] * U_d[u[0]] + nv[3 * f + 1] * U_d[u[1]] + nv[3 * f + 2] * U_d[u[2]];
            t += W[16 * (size_t)f + a * 4 + b] * un;
        }
        acc += t;
    }
    F_p[n] = acc;
}

void boundary_integral3(Ctx& c, const double* U_d, double* F_p) {
    const int nn = c.g3.nn;
    if (!c.fb_off) {
        launch_fill(c, F_p, 0.0, nn);
        return;
    }
    k_boundary_integral3<<<(nn + 255) / 256, 256, 0, c.stream>>>(nn, c.fb_off, c.fb_edge, c.fb_w, c.fb_geom,
                                                                  c.fb_ui, U_d, F_p);
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ mass matvec
// y = M_c x per component (the 3-component consistent_mass_mult, element_mult.py:11-36)
struct Mass3Body {
    Grid3D g; const ElemConst3* ec; const double* x; double* y;
    __device__ void operator()(int n, double (&)[1]) const {
        Nbr3 s;
        node_nbr3(g, n, s);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double X[27];
            const double* xc = x + (size_t)d * g.nn;
            load27(s, [&](int k) { return xc[k]; }, X);
            y[n + (size_t)d * g.nn] = ebe3(s, ec->Me, X);
        }
    }
};
void mass_matvec3(Ctx& c, const double* x, double* y) {
    launch_node<0>(c, c.g3.nn, nullptr, Mass3Body{c.g3, c.ec3_dev, x, y}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ the step
void four_step3(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* U_d,
                const double* F_p, const double* F, double* U, double* P, int* it) {
    scatter_ud(c, U_d);
    build_ce3(c, U_n, 0.5);
    it[0] = solve_momentum3(c, U_n, P_n, rho, mu, dt, F, c.m_x);
    launch_stencil3<0>(c, c.g3, nullptr,
                       Proj3S{c.g3, c.ec3_dev, c.m_x, P_n, c.Mlump, nullptr, nullptr, c.ustar, dt / rho}, NoFin{});
    FCFD_CHECK_LAUNCH();
    it[1] = solve_pressure3(c, c.ustar, P_n, F_p, rho, dt, P);
    launch_stencil3<0>(c, c.g3, nullptr,
                       Proj3S{c.g3, c.ec3_dev, c.ustar, P, c.Mlump, c.ud_full, c.dmask, U, -dt / rho}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

}  // namespace fcfd
