// Geometric multigrid V-cycle on Q1 hex8 box grids (FP64), the 3D pressure preconditioner.
//
// The reference stops at 2D (simulation.py:75-76: "3D multigrid not yet implemented").  This is
// the V-cycle of multigrid.py:231-330 carried over to box grids unchanged in structure:
//   - the x ratio sets the level count and all three directions halve in lockstep (:12, :80);
//   - each level rediscretises Le on its own element (in 3D Le scales with h);
//   - down: x = 2/3 r / L_diag (one Jacobi from zero), f = r - L x, restriction R = P^T;
//   - coarse: CG to tol with the pinned rows zeroed (pinned form) or one Jacobi from zero
//     (Neumann form);
//   - up: x += P y, one damped Jacobi (omega = 2/3) against the level's residual;
//   - the Neumann form mean-projects as the 2D V-cycle does, with the means applied lazily
//     through the same device-state slots.
// P is trilinear interpolation applied along x, then y, then z; R = P^T is applied in the same
// order (c[i] = f[2i] + 0.5 (f[2i-1] + f[2i+1]) per direction), so both are bitwise the
// separable transfers of the oracle.
#include "fcfd_internal.hpp"
#include "fcfd_hex.hpp"
#include "fcfd_cg.hpp"
#include "fcfd_mgnodes.hpp"
#include "fcfd_mg3nodes.hpp"

namespace fcfd {

// ------------------------------------------------------------------------------ down leg
// f = r - L x: pinned (x = 2/3 r / L_diag with the pins 0, pinned rows of L x 0) or Neumann (x
// and r mean-projected)
struct DownRes3 {
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; double* x; double* f;
    const DevState* st; int l; int neumann;
    __device__ void operator()(int n, double (&acc)[1]) const {
        Nbr3 s;
        node_nbr3(g, n, s);
        double X[27];
        double fv;
        if (neumann) {
            const double mx = st->mg[l][MG_X], mr = st->mg[l][MG_R];
            load27(s, [&](int k) { return x[k] - mx; }, X);
            fv = (r[n] - mr) - ebe3(s, Le, X);
        } else {
            load27(s, [&](int k) {
                const double v = (2.0 / 3.0 * r[k]) / Ld[k];
                return (pin && pin[k]) ? 0.0 : v;
            }, X);
            x[n] = X[S3C];
            const bool pn = pin && pin[n];
            const double rh = pn ? 0.0 : r[n];
            const double lx = pn ? 0.0 : ebe3(s, Le, X);
            fv = rh - lx;
        }
        f[n] = fv;
        acc[0] += fv;
    }
};

// x <- x + P yc; one damped Jacobi (multigrid.py:214-228, :306-326)
struct Jacobi3 {
    Grid3D gf, gc; const double* Leoff; const double* Ld; const uint8_t* pin; const double* r; const double* x;
    const double* p; const double* yc; double* y; const DevState* st; int l; int neumann;
    __device__ void operator()(int n, double (&acc)[1]) const {
        Nbr3 s;
        node_nbr3(gf, n, s);
        double X[27];
        double rh;
        if (neumann) {
            const double mx = st->mg[l][MG_X], mp = st->mg[l][MG_P];
            load27(s, [&](int k) { return (x[k] - mx) + (p[k] - mp); }, X);
            const double oup = l == 0 ? 0.0 : st->mg[l][MG_R];
            rh = r[n] - oup;
        } else {
            const int k = n / gf.nxy, rem = n - k * gf.nxy;
            const int j = rem / gf.nnx, i = rem - j * gf.nnx;
            double Y[27];
            interp27(gc, yc, i, j, k, Y);
            load27(s, [&](int m) { return x[m]; }, X);
#pragma unroll
            for (int t = 0; t < 27; ++t) {
                const int m = s.nb[t];
                const bool ok = m != n || t == S3C;  // clamped ids of missing neighbours point at n
                const bool keep = ok && !(pin && pin[m]);
                X[t] = keep ? X[t] + Y[t] : 0.0;
            }
            rh = r[n];
        }
        const double off = ebe3(s, Leoff, X);
        double v = (OMEGA * (rh - off)) / Ld[n] + (1.0 - OMEGA) * X[S3C];
        if (!neumann && pin && pin[n]) v = 0.0;
        y[n] = v;
        acc[0] += v;
    }
};

// LDS-staged forms of the pinned DownRes3 / Jacobi3 (stencil3t_kernel): the pre-smoothed x, and
// x + P yc, are formed once per window entry with the node kernels' operands (interp3_at is
// bitwise interp27), and every node reads its 27 values from LDS.
struct DownRes3S {
    static constexpr int NV = 1;
    Grid3D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; double* x; double* f;
    __device__ __forceinline__ void stage(int id, double (&v)[1]) const {
        const double xv = (2.0 / 3.0 * r[id]) / Ld[id];
        v[0] = pin[id] ? 0.0 : xv;
    }
    template <class SV>
    __device__ void operator()(int n, const SV& sv, double (&)[1]) const {
        Nbr3 s;
        node_elems3(g, n, s);
        double X[27];
        view27(sv, 0, n, X);
        x[n] = X[S3C];
        const bool pn = pin[n];
        const double rh = pn ? 0.0 : r[n];
        const double lx = pn ? 0.0 : ebe3(s, Le, X);
        f[n] = rh - lx;
    }
};
struct Jacobi3S {
    static constexpr int NV = 1;
    Grid3D gf, gc; const double* Leoff; const double* Ld; const uint8_t* pin; const double* r; const double* x;
    const double* yc; double* y;
    __device__ __forceinline__ void stage(int id, double (&v)[1]) const {
        const double xv = x[id] + interp3_at(gf, gc, yc, 0.0, id);
        v[0] = pin[id] ? 0.0 : xv;
    }
    template <class SV>
    __device__ void operator()(int n, const SV& sv, double (&)[1]) const {
        Nbr3 s;
        node_elems3(gf, n, s);
        double X[27];
        view27(sv, 0, n, X);
        const double off = ebe3(s, Leoff, X);
        double v = (OMEGA * (r[n] - off)) / Ld[n] + (1.0 - OMEGA) * X[S3C];
        if (pin[n]) v = 0.0;
        y[n] = v;
    }
};
template <class Body>
static void launch_stencil3_nored(Ctx& c, const Grid3D& g, const int* active, const Body& body) {
    const dim3 t = s3t_tiles(g);
    stencil3t_kernel<0><<<t.x * t.y * t.z, BLOCK, 0, c.stream>>>(g, (int)t.x, (int)t.y, active, body, c.partials,
                                                                 c.pstride);
}

// ------------------------------------------------------------------------------ coarse level
// Pinned: CG (M = identity) to tol, maxiter 1000, on L with the pinned rows zeroed
// (multigrid.py:278-287), in one workgroup; the search direction lives in LDS, x / r / Ap in
// memory (L2-resident).
__global__ __launch_bounds__(1024) void k_coarse_cg3(Grid3D g, const ElemConst3* __restrict__ ec,
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
            Nbr3 s;
            node_nbr3(g, n, s);
            double X[27];
            load27(s, [&](int m) { return p[m]; }, X);
            const double a = (pin && pin[n]) ? 0.0 : ebe3(s, ec->Le, X);
            cAp[n] = a;
            pap[0] += X[S3C] * a;
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

// ------------------------------------------------------------------------------ set-up
void build_mg3(Ctx& c, const fcfd_solver_desc* s) {
    const int nlev = s->mg_levels;
    FCFD_REQUIRE(nlev >= 1 && nlev <= MAX_LEVELS, "bad multigrid level count");
    const Grid3D& g0 = c.g3;
    FCFD_REQUIRE((g0.nex >> (nlev - 1)) << (nlev - 1) == g0.nex && (g0.ney >> (nlev - 1)) << (nlev - 1) == g0.ney &&
                     (g0.nez >> (nlev - 1)) << (nlev - 1) == g0.nez && (g0.nex >> (nlev - 1)) >= 1 &&
                     (g0.ney >> (nlev - 1)) >= 1 && (g0.nez >> (nlev - 1)) >= 1,
                 "3D multigrid: every element count must halve exactly at each level");
    c.nlev = nlev;
    c.lv.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
        Level& L = c.lv[l];
        L.g3 = make_grid3(g0.nex >> l, g0.ney >> l, g0.nez >> l);
        L.g = Grid2D{0, 0, 0, 0, L.g3.nn, L.g3.ne, 0};
        ElemConst3 ec;
        hex8_constants(c.lx / L.g3.nex, c.ly / L.g3.ney, c.lz / L.g3.nez, ec);
        L.ec3_dev = dalloc<ElemConst3>(c, 1);
        FCFD_HIP(hipMemcpy(L.ec3_dev, &ec, sizeof(ElemConst3), hipMemcpyHostToDevice));
        double dg[8];
        for (int a = 0; a < 8; ++a) dg[a] = ec.Le[a * 8 + a];
        std::vector<double> Ld;
        node_sums3(L.g3, dg, Ld);
        if (!L.Ld) L.Ld = dalloc<double>(c, L.g3.nn);
        FCFD_HIP(hipMemcpy(L.Ld, Ld.data(), sizeof(double) * Ld.size(), hipMemcpyHostToDevice));
        const int n = L.g3.nn;
        if (l > 0) L.r = dalloc<double>(c, n);
        if (!L.x) L.x = dalloc<double>(c, n);
        if (!L.y) L.y = dalloc<double>(c, n);
        L.f = dalloc<double>(c, n);
        if (c.set_zeroP) {
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
            FCFD_HIP(hipMemcpy(L.pin, pin.data(), n, hipMemcpyHostToDevice));
        }
    }
    const int ncoarse = c.lv[nlev - 1].g3.nn;
    FCFD_REQUIRE(!c.set_zeroP || (size_t)ncoarse * sizeof(double) <= 64 * 1024,
                 "3D multigrid: the pinned coarse grid must have at most 8192 nodes");
    c.cc_r = dalloc<double>(c, ncoarse);
    c.cc_Ap = dalloc<double>(c, ncoarse);
    // mixed form, pinned pressure: the coarse solve is the exact inverse of the pinned coarse
    // operator, stored FP32 (the FP64 form keeps CG to tol)
    if (c.precision == FCFD_PREC_MIXED && c.set_zeroP && nlev > 1 && ncoarse <= COARSE3_DENSE_MAX) {
        const Grid3D& g = c.lv[nlev - 1].g3;
        ElemConst3 ec;
        hex8_constants(c.lx / g.nex, c.ly / g.ney, c.lz / g.nez, ec);
        std::vector<double> A((size_t)ncoarse * ncoarse, 0.0), inv;
        for (int ke = 0; ke < g.nez; ++ke)
            for (int je = 0; je < g.ney; ++je)
                for (int ie = 0; ie < g.nex; ++ie) {
                    const int n0 = ke * g.nxy + je * g.nnx + ie;
                    for (int a = 0; a < 8; ++a)
                        for (int b = 0; b < 8; ++b)
                            A[(size_t)(n0 + hbz(a) * g.nxy + hby(a) * g.nnx + hbx(a)) * ncoarse + n0 +
                              hbz(b) * g.nxy + hby(b) * g.nnx + hbx(b)] += ec.Le[a * 8 + b];
                }
        std::vector<uint8_t> pin(ncoarse);
        FCFD_HIP(hipMemcpy(pin.data(), c.lv[nlev - 1].pin, ncoarse, hipMemcpyDeviceToHost));
        std::vector<int64_t> pins;
        for (int k = 0; k < ncoarse; ++k)
            if (pin[k]) pins.push_back(k);
        pin_identity(A, ncoarse, pins);
        if (spd_inverse_banded(A, ncoarse, g.nxy + g.nnx + 1, inv, true)) {
            std::vector<float> inv32(inv.begin(), inv.end());
            c.coarse_inv32 = dalloc<float>(c, inv32.size());
            FCFD_HIP(hipMemcpy(c.coarse_inv32, inv32.data(), sizeof(float) * inv32.size(), hipMemcpyHostToDevice));
        }
    }
}

// ------------------------------------------------------------------------------ V-cycle driver
// r0: the fine residual (its mean in mg[0][MG_R] for the Neumann form); result in lv[0].y (its
// mean in mg[0][MG_Y] for the Neumann form, applied by the consumer)
const double* vcycle3(Ctx& c, const double* r0, const int* active) {
    const int N = c.nlev;
    const int neu = !c.set_zeroP;
    const int ld = N - 1;
    auto run = [&](int n, const auto& body, const MeanFin& fin) {
        if (neu) launch_node<1>(c, n, active, body, fin);
        else launch_node<0>(c, n, active, body, fin);
    };
    auto rl = [&](int l) -> const double* { return l == 0 ? r0 : c.lv[l].r; };
    for (int l = 0; l < ld; ++l) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        const int n = L.g3.nn;
        if (neu)
            launch_node<1>(c, n, active, DownX<double, double>{rl(l), L.Ld, L.x, c.st, l},
                           MeanFin{&c.st->mg[l][MG_X], nullptr, n});
        if (neu)
            run(n, DownRes3{L.g3, L.ec3_dev->Le, rl(l), L.Ld, L.pin, L.x, L.f, c.st, l, neu},
                MeanFin{&c.st->mg[l][MG_F], nullptr, n});
        else
            launch_stencil3_nored(c, L.g3, active, DownRes3S{L.g3, L.ec3_dev->Le, rl(l), L.Ld, L.pin, L.x, L.f});
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
            k_coarse_cg3<<<1, 1024, shm, c.stream>>>(C.g3, C.ec3_dev, C.pin, rl(ld), C.y, c.cc_r, c.cc_Ap, c.st,
                                                     c.tol, 1000, active);
        }
        FCFD_CHECK_LAUNCH();
    }
    for (int l = ld - 1; l >= 0; --l) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        const int n = L.g3.nn;
        if (neu)
            launch_node<1>(c, n, active, Prolong3{L.g3, Lc.g3, Lc.y, L.f, c.st, l, 1},
                           MeanFin{&c.st->mg[l][MG_P], nullptr, n});
        if (neu)
            run(n, Jacobi3{L.g3, Lc.g3, L.ec3_dev->Leoff, L.Ld, L.pin, rl(l), L.x, L.f, Lc.y, L.y, c.st, l, neu},
                MeanFin{&c.st->mg[l][MG_Y], nullptr, n});
        else
            launch_stencil3_nored(c, L.g3, active,
                                  Jacobi3S{L.g3, Lc.g3, L.ec3_dev->Leoff, L.Ld, L.pin, rl(l), L.x, Lc.y, L.y});
    }
    FCFD_CHECK_LAUNCH();
    return c.lv[0].y;
}

// z = y0 - mean (Neumann)
struct ZOut3 {
    const double* y; double* z; const DevState* st; int neumann;
    __device__ void operator()(int k, double (&)[1]) const { z[k] = y[k] - (neumann ? st->mg[0][MG_Y] : 0.0); }
};

void vcycle3_api(Ctx& c, const double* r, double* z) {
    FCFD_REQUIRE(c.precond == FCFD_PC_MULTIGRID, "context was not created with the multigrid preconditioner");
    const int nn = c.g3.nn;
    struct SumBody {
        const double* r;
        __device__ void operator()(int k, double (&acc)[1]) const { acc[0] += r[k]; }
    };
    launch_node<1>(c, nn, nullptr, SumBody{r}, MeanFin{&c.st->mg[0][MG_R], nullptr, nn});
    const double* y = vcycle3(c, r, nullptr);
    launch_node<0>(c, nn, nullptr, ZOut3{y, z, c.st, !c.set_zeroP}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

}  // namespace fcfd
