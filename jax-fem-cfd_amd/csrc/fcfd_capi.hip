// Context construction and the extern "C" entry points declared in include/fcfd.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <string>
#include "fcfd_internal.hpp"

using namespace fcfd;

namespace fcfd {
void enter(Ctx& c) {
    FCFD_HIP(hipEventRecord(c.ev_in, c.user_stream));
    FCFD_HIP(hipStreamWaitEvent(c.stream, c.ev_in, 0));
}
void leave(Ctx& c) {
    FCFD_HIP(hipEventRecord(c.ev_out, c.stream));
    FCFD_HIP(hipStreamWaitEvent(c.user_stream, c.ev_out, 0));
}
void graph_reset(IterGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    for (int p = 0; p < 2; ++p) {
        if (g.one[p]) (void)hipGraphExecDestroy(g.one[p]);
        g.one[p] = nullptr;
    }
}
}  // namespace fcfd

struct fcfd_ctx {
    Ctx c;
    double* scratch = nullptr;  // address of `scratch` on this context's GPU
};

// Block partials of fcfd_dot_noalloc.  Each GPU the library runs on holds its own copy until the
// process ends.  Host code never passes the array's name to a kernel (in host code it is only a
// stand-in): it uses the address that hipGetSymbolAddress returns on the current device.
static __device__ __align__(256) double scratch[FCFD_DOT_NOALLOC_MAX_N / BLOCK];
static_assert(FCFD_DOT_NOALLOC_MAX_N % BLOCK == 0, "the partials of the largest n fill the array exactly");

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Address of `scratch` on the current device, looked up once per device.
double* scratch_on_current_device() {
    int dev = 0;
    FCFD_HIP(hipGetDevice(&dev));
    FCFD_REQUIRE(dev >= 0 && dev < kMaxDevices, "device id beyond the scratch address table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        FCFD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    return g_scratch_dev[dev];
}
}  // namespace

static thread_local std::string g_last_error;

template <class F>
static int guard(F&& f) {
    try {
        f();
        return FCFD_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return FCFD_ERR_ARG;
    }
}

// guard + stream join: the call's work runs on the context's private stream, ordered after
// everything already queued on the caller's stream, and the caller's stream waits for it.
template <class F>
static int run(fcfd_ctx* ctx, F&& f) {
    try {
        enter(ctx->c);
        f();
        leave(ctx->c);
        return FCFD_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return FCFD_ERR_ARG;
    }
}

// entry points built for 2D contexts only
template <class F>
static int run2d(fcfd_ctx* ctx, const char* what, F&& f) {
    if (ctx->c.dim != 2 || ctx->c.order != 1) {
        g_last_error = std::string(what) + " is available for 2D Q1 contexts only";
        return FCFD_ERR_UNSUPPORTED;
    }
    return run(ctx, std::forward<F>(f));
}

namespace fcfd {

// Inverse of an SPD matrix A (n x n, row-major, entries beyond half-bandwidth w from the diagonal
// zero): banded Cholesky A = R^T R in place, then one forward / back substitution per unit
// vector.  Returns false if A is not numerically SPD.  With `symmetrise`, inv is replaced by
// (X + X^T) / 2, so that its FP32 rounding is symmetric too.
bool spd_inverse_banded(std::vector<double>& A, int n, int w, std::vector<double>& inv, bool symmetrise) {
    auto lo = [&](int i) { return i - w > 0 ? i - w : 0; };
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int k = lo(j); k < j; ++k) d -= A[(size_t)k * n + j] * A[(size_t)k * n + j];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        A[(size_t)j * n + j] = d;
        const int ie = j + w < n ? j + w + 1 : n;
        for (int i = j + 1; i < ie; ++i) {
            double v = A[(size_t)j * n + i];
            for (int k = lo(i); k < j; ++k) v -= A[(size_t)k * n + j] * A[(size_t)k * n + i];
            A[(size_t)j * n + i] = v / d;
        }
    }
    inv.assign((size_t)n * n, 0.0);
    std::vector<double> z(n);
    for (int c = 0; c < n; ++c) {
        // R^T z = e_c (z = 0 above c), then R x = z
        for (int i = 0; i < c; ++i) z[i] = 0.0;
        for (int i = c; i < n; ++i) {
            double v = i == c ? 1.0 : 0.0;
            for (int k = lo(i) > c ? lo(i) : c; k < i; ++k) v -= A[(size_t)k * n + i] * z[k];
            z[i] = v / A[(size_t)i * n + i];
        }
        double* x = &inv[(size_t)c * n];
        for (int i = n - 1; i >= 0; --i) {
            double v = z[i];
            const int ke = i + w < n ? i + w + 1 : n;
            for (int k = i + 1; k < ke; ++k) v -= A[(size_t)i * n + k] * x[k];
            x[i] = v / A[(size_t)i * n + i];
        }
    }
    for (int i = 0; i < n && symmetrise; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double v = 0.5 * (inv[(size_t)i * n + j] + inv[(size_t)j * n + i]);
            inv[(size_t)i * n + j] = inv[(size_t)j * n + i] = v;
        }
    return true;
}

// Replace the rows and columns of the pinned nodes by those of the identity.
void pin_identity(std::vector<double>& A, int n, const std::vector<int64_t>& pins) {
    for (int64_t p : pins) {
        for (int k = 0; k < n; ++k) A[(size_t)p * n + k] = A[(size_t)k * n + p] = 0.0;
        A[(size_t)p * n + p] = 1.0;
    }
}

}  // namespace fcfd

namespace {

// L_diag and lumped mass on a level: per node, the sum over its incident elements (in element
// order) of diag(Le)[a] and of the row sums of Me (element_calc.py:9-25, :99-109).
void node_sums(const Grid2D& g, const double* per_row, std::vector<double>& out) {
    out.assign(g.nn, 0.0);
    for (int e = 0; e < g.ne; ++e) {
        int je = e / g.nex, ie = e - je * g.nex;
        int c[4];
        if (g.periodic) {
            int ip = (ie + 1) % g.nnx, jp = (je + 1) % g.nny;
            c[0] = je * g.nnx + ie; c[1] = je * g.nnx + ip; c[2] = jp * g.nnx + ip; c[3] = jp * g.nnx + ie;
        } else {
            int n0 = je * g.nnx + ie;
            c[0] = n0; c[1] = n0 + 1; c[2] = n0 + g.nnx + 1; c[3] = n0 + g.nnx;
        }
        for (int a = 0; a < 4; ++a) out[c[a]] += per_row[a];
    }
}

template <class T>
void upload(const std::vector<T>& h, T* d) {
    FCFD_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

// Dense inverse of a level's Laplacian with its pinned rows and columns replaced by the identity
// (the operator the pinned coarse CG inverts: x = 0 at the pins), by Cholesky in FP64.  Returns
// false if the matrix is not positive definite (no pin on a pure-Neumann level).  Node j * nnx + i
// couples only to nodes within nnx + 1 of it on a non-periodic grid, so the factor R (A = R^T R)
// lives in that band and every loop below skips the terms outside it (they are exact zeros).
constexpr int COARSE_DENSE_MAX = 1089;  // 33 x 33 nodes: 4.7 MB of FP32
bool coarse_inverse(const Level& L, std::vector<double>& inv) {
    const Grid2D& g = L.g;
    const int n = g.nn;
    const int w = g.periodic ? n : g.nnx + 1;  // half-bandwidth
    std::vector<double> A((size_t)n * n, 0.0);
    for (int e = 0; e < g.ne; ++e) {
        const int je = e / g.nex, ie = e - je * g.nex;
        const int ip = g.periodic && ie + 1 == g.nnx ? 0 : ie + 1;
        const int jp = g.periodic && je + 1 == g.nny ? 0 : je + 1;
        const int cn[4] = {je * g.nnx + ie, je * g.nnx + ip, jp * g.nnx + ip, jp * g.nnx + ie};
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) A[(size_t)cn[a] * n + cn[b]] += L.ec.Le[a * 4 + b];
    }
    pin_identity(A, n, L.pins);
    return spd_inverse_banded(A, n, w, inv, false);
}

void build_ctx(Ctx& c, const fcfd_mesh_desc* m, const fcfd_solver_desc* s, void* stream) {
    FCFD_REQUIRE(m && s, "null descriptor");
    FCFD_REQUIRE(m->ndim == 2 || m->ndim == 3, "ndim must be 2 or 3");
    FCFD_REQUIRE(m->nex >= 1 && m->ney >= 1, "bad element counts");
    c.user_stream = (hipStream_t)stream;
    FCFD_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    FCFD_HIP(hipEventCreateWithFlags(&c.ev_in, hipEventDisableTiming));
    FCFD_HIP(hipEventCreateWithFlags(&c.ev_out, hipEventDisableTiming));
    for (auto& e : c.ev_spec) FCFD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    {
        // solver chunks as captured HIP graphs (FCFD_GRAPHS=1) or as direct launches (default:
        // measured faster at 1024^2, where each graph launch costs more device time than the
        // launches it replaces)
        const char* e = std::getenv("FCFD_GRAPHS");
        c.use_graphs = e && e[0] == '1';
        const char* t = std::getenv("FCFD_MG_TILED");
        c.mg_tiled = !(t && t[0] == '0');
        // a fixed chunk length for the solves of speculative steps (default: predicted)
        const char* k = std::getenv("FCFD_SPEC_CHUNK");
        c.spec_chunk_fixed = k ? std::max(0, std::atoi(k)) : 0;
    }
    c.tol = s->tol;
    c.set_zeroP = s->set_zeroP;
    c.precond = s->precond;
    c.precision = s->precision;
    c.ce_format = s->ce_format;
    c.maxit_mom = s->maxiter_momentum;
    c.maxit_p = s->maxiter_pressure;
    c.approx_keep = s->approx_inv_variant == 1;
    FCFD_REQUIRE(c.precond == FCFD_PC_NONE || c.precond == FCFD_PC_JACOBI || c.precond == FCFD_PC_APPROX_INV ||
                     c.precond == FCFD_PC_MULTIGRID,
                 "unknown pressure preconditioner");
    if (m->ndim == 3) {
        FCFD_REQUIRE(m->order == 0 || m->order == 1 || m->order == 2, "order must be 1 (hex8) or 2 (hex27)");
        if (m->order == 2) build_ctx_h27(c, m, s);
        else build_ctx3(c, m, s);
        return;
    }
    FCFD_REQUIRE(m->order == 0 || m->order == 1 || m->order == 2, "order must be 1 (quad4) or 2 (quad9)");
    if (m->order == 2) {
        build_ctx_q2(c, m, s);
        return;
    }
    c.g = make_grid(m->nex, m->ney, m->periodic);
    c.lx = m->lx; c.ly = m->ly;
    const Grid2D& g = c.g;
    const int nn = g.nn;
    quad4_constants(c.lx / g.nex, c.ly / g.ney, c.ec);
    c.ec_dev = dalloc<ElemConst>(c, 1);
    FCFD_HIP(hipMemcpy(c.ec_dev, &c.ec, sizeof(ElemConst), hipMemcpyHostToDevice));

    // lumped mass, tiled per component
    {
        double rs[4];
        for (int a = 0; a < 4; ++a) {
            rs[a] = 0.0;
            for (int b = 0; b < 4; ++b) rs[a] += c.ec.Me[a * 4 + b];
        }
        std::vector<double> M;
        node_sums(g, rs, M);
        M.resize(2 * nn);
        std::copy(M.begin(), M.begin() + nn, M.begin() + nn);
        c.Mlump = dalloc<double>(c, 2 * nn);
        upload(M, c.Mlump);
    }

    // multigrid levels (level 0 always exists: L_diag and scratch of the fine grid)
    int nlev = c.precond == FCFD_PC_MULTIGRID ? s->mg_levels : 1;
    FCFD_REQUIRE(nlev >= 1 && nlev <= MAX_LEVELS, "bad number of multigrid levels");
    if (nlev > 1) {
        FCFD_REQUIRE((g.nex >> (nlev - 1)) >= 1 && (g.ney >> (nlev - 1)) >= 1, "too many multigrid levels");
        FCFD_REQUIRE(g.nex % (1 << (nlev - 1)) == 0 && g.ney % (1 << (nlev - 1)) == 0,
                     "element counts must be divisible by 2^(levels-1)");
    }
    c.nlev = nlev;
    c.lv.resize(nlev);
    const bool mixed = c.precision == FCFD_PREC_MIXED;
    for (int l = 0; l < nlev; ++l) {
        Level& L = c.lv[l];
        L.g = make_grid(g.nex >> l, g.ney >> l, g.periodic);
        quad4_constants(c.lx / L.g.nex, c.ly / L.g.ney, L.ec);
        L.ec_dev = dalloc<ElemConst>(c, 1);
        FCFD_HIP(hipMemcpy(L.ec_dev, &L.ec, sizeof(ElemConst), hipMemcpyHostToDevice));
        double dg[4] = {L.ec.Le[0], L.ec.Le[5], L.ec.Le[10], L.ec.Le[15]};
        std::vector<double> Ld;
        node_sums(L.g, dg, Ld);
        L.Ld = dalloc<double>(c, L.g.nn);
        upload(Ld, L.Ld);
        std::vector<float> Ld32(Ld.begin(), Ld.end());
        L.Ld32 = dalloc<float>(c, L.g.nn);
        upload(Ld32, L.Ld32);
        // L_diag by node class (first / interior / last along each direction); the tiled V-cycle
        // kernels read it from this table, so every node must carry its class's value
        {
            auto rep = [](int cls, int n) { return cls == 0 ? 0 : (cls == 2 ? n - 1 : (n > 2 ? 1 : 0)); };
            for (int cj = 0; cj < 3; ++cj)
                for (int ci = 0; ci < 3; ++ci)
                    L.ldc[cj][ci] = Ld[(size_t)rep(cj, L.g.nny) * L.g.nnx + rep(ci, L.g.nnx)];
            for (int j = 0; j < L.g.nny && c.mg_tiled; ++j)
                for (int i = 0; i < L.g.nnx; ++i) {
                    const int ci = i == 0 ? 0 : (i == L.g.nnx - 1 ? 2 : 1);
                    const int cj = j == 0 ? 0 : (j == L.g.nny - 1 ? 2 : 1);
                    if (Ld[(size_t)j * L.g.nnx + i] != L.ldc[cj][ci]) {
                        c.mg_tiled = false;
                        break;
                    }
                }
        }
        const int n = L.g.nn;
        if (l > 0) L.r = dalloc<double>(c, n);
        L.x = dalloc<double>(c, n);
        L.y = dalloc<double>(c, n);
        L.f = dalloc<double>(c, n);
        if (mixed) {
            if (l > 0) L.r32 = dalloc<float>(c, n);
            L.x32 = dalloc<float>(c, n);
            L.y32 = dalloc<float>(c, n);
            L.f32 = dalloc<float>(c, n);
        }
        if (c.set_zeroP) {
            std::vector<uint8_t> pin(n, 0);
            if (s->mg_gamma_p && s->mg_gamma_p_off) {
                for (int64_t k = s->mg_gamma_p_off[l]; k < s->mg_gamma_p_off[l + 1]; ++k) {
                    int64_t v = s->mg_gamma_p[k];
                    FCFD_REQUIRE(v >= 0 && v < n, "multigrid pin out of range");
                    pin[v] = 1;
                    L.pins.push_back(v);
                }
            } else {
                pin[0] = 1;
                L.pins.push_back(0);
            }
            L.pin = dalloc<uint8_t>(c, n);
            upload(pin, L.pin);
        }
    }
    {
        const int ncoarse = c.lv[nlev - 1].g.nn;
        c.cc_r = dalloc<double>(c, ncoarse);
        c.cc_p = dalloc<double>(c, ncoarse);
        c.cc_Ap = dalloc<double>(c, ncoarse);
        if (mixed) {
            c.cc_r32 = dalloc<float>(c, ncoarse);
            c.cc_p32 = dalloc<float>(c, ncoarse);
            c.cc_Ap32 = dalloc<float>(c, ncoarse);
        }
        // mixed form, pinned pressure: the coarse solve is the exact inverse of the pinned
        // coarse operator (the FP64 form keeps the reference's CG to tol)
        if (mixed && c.set_zeroP && nlev > 1 && ncoarse <= COARSE_DENSE_MAX) {
            std::vector<double> inv;
            if (coarse_inverse(c.lv[nlev - 1], inv)) {
                std::vector<float> inv32(inv.begin(), inv.end());
                c.coarse_inv32 = dalloc<float>(c, inv32.size());
                upload(inv32, c.coarse_inv32);
            }
        }
    }

    // fine-level pressure pins (the approximate-inverse preconditioner zeroes the gamma_p rows of
    // its two inner products in both forms, standard.py:225-234)
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
        upload(pin, c.ppin);
    }

    // Dirichlet DOFs
    std::vector<int64_t> gd;
    if (s->gamma_d && s->n_gamma_d > 0) {
        gd.assign(s->gamma_d, s->gamma_d + s->n_gamma_d);
        for (size_t k = 0; k < gd.size(); ++k) {
            FCFD_REQUIRE(gd[k] >= 0 && gd[k] < 2 * nn, "gamma_d out of range");
            FCFD_REQUIRE(k == 0 || gd[k] > gd[k - 1], "gamma_d must be sorted and unique");
        }
        c.n_gamma_d = (int)gd.size();
        std::vector<uint8_t> mask(2 * nn, 0);
        for (auto v : gd) mask[v] = 1;
        c.dmask = dalloc<uint8_t>(c, 2 * nn);
        upload(mask, c.dmask);
        c.gamma_d_dev = dalloc<int64_t>(c, gd.size());
        upload(gd, c.gamma_d_dev);
        c.ud_full = dalloc<double>(c, 2 * nn);
        FCFD_HIP(hipMemset(c.ud_full, 0, sizeof(double) * 2 * nn));
    }

    // passive-scalar Dirichlet nodes
    if (s->gamma_d_phi && s->n_gamma_d_phi > 0) {
        std::vector<int64_t> gp(s->gamma_d_phi, s->gamma_d_phi + s->n_gamma_d_phi);
        std::vector<uint8_t> mask(nn, 0);
        for (auto v : gp) {
            FCFD_REQUIRE(v >= 0 && v < nn, "gamma_d_phi out of range");
            mask[v] = 1;
        }
        c.n_gamma_d_phi = (int)gp.size();
        c.gamma_d_phi_dev = dalloc<int64_t>(c, gp.size());
        upload(gp, c.gamma_d_phi_dev);
        c.phi_mask = dalloc<uint8_t>(c, nn);
        upload(mask, c.phi_mask);
        c.phi_ud = dalloc<double>(c, nn);
        FCFD_HIP(hipMemset(c.phi_ud, 0, sizeof(double) * nn));
    }

    // boundary edges -> per-node gather lists in bconn order (segment_sum order)
    if (s->bconn && s->nvec && s->n_bedge > 0 && !gd.empty()) {
        const int ne_b = (int)s->n_bedge;
        c.n_bedge = ne_b;
        std::vector<std::vector<int>> lists(nn);
        std::vector<double> geom(3 * ne_b);
        std::vector<int> ui(4 * ne_b);
        const double dx = c.lx / g.nex, dy = c.ly / g.ney;
        auto sorted_index = [&](int64_t v) {
            return (int)(std::lower_bound(gd.begin(), gd.end(), v) - gd.begin());
        };
        for (int ed = 0; ed < ne_b; ++ed) {
            int64_t a = s->bconn[2 * ed], b = s->bconn[2 * ed + 1];
            FCFD_REQUIRE(a >= 0 && a < nn && b >= 0 && b < nn, "bconn out of range");
            double xa = (double)(a % g.nnx) * dx, ya = (double)(a / g.nnx) * dy;
            double xb = (double)(b % g.nnx) * dx, yb = (double)(b / g.nnx) * dy;
            geom[3 * ed + 0] = (double)s->nvec[2 * ed];
            geom[3 * ed + 1] = (double)s->nvec[2 * ed + 1];
            geom[3 * ed + 2] = std::sqrt((xb - xa) * (xb - xa) + (yb - ya) * (yb - ya));
            int ia[4] = {sorted_index(a), sorted_index(a + nn), sorted_index(b), sorted_index(b + nn)};
            for (int k = 0; k < 4; ++k) {
                FCFD_REQUIRE(ia[k] < (int)gd.size(), "boundary node is not a Dirichlet DOF");
                ui[4 * ed + k] = ia[k];
            }
            lists[a].push_back(ed * 2 + 0);
            lists[b].push_back(ed * 2 + 1);
        }
        std::vector<int> off(nn + 1, 0), slots;
        for (int n = 0; n < nn; ++n) {
            off[n + 1] = off[n] + (int)lists[n].size();
            slots.insert(slots.end(), lists[n].begin(), lists[n].end());
        }
        c.fb_off = dalloc<int>(c, nn + 1);
        upload(off, c.fb_off);
        c.fb_edge = dalloc<int>(c, slots.size());
        upload(slots, c.fb_edge);
        c.fb_geom = dalloc<double>(c, geom.size());
        upload(geom, c.fb_geom);
        c.fb_ui = dalloc<int>(c, ui.size());
        upload(ui, c.fb_ui);
    }

    // element matrix storage for the timestep
    {
        const size_t ne = g.ne;
        switch (c.ce_format) {
            case FCFD_CE_F64: c.ce = dalloc<double>(c, 16 * ne); break;
            case FCFD_CE_F32: c.ce = dalloc<float>(c, 16 * ne); break;
            case FCFD_CE_F16S:
            case FCFD_CE_BF16S:
                c.ce = dalloc<uint16_t>(c, 16 * ne);
                c.ce_scale = dalloc<float>(c, ne);
                break;
            default: throw Error(FCFD_ERR_ARG, "unknown Ce format");
        }
    }
    // solver work
    c.m_r = dalloc<double>(c, 2 * nn);
    c.m_rhat = dalloc<double>(c, 2 * nn);
    for (int k = 0; k < 2; ++k) {
        c.m_p[k] = dalloc<double>(c, 2 * nn);
        c.m_q[k] = dalloc<double>(c, 2 * nn);
        c.c_p[k] = dalloc<double>(c, nn);
        c.c_z[k] = dalloc<double>(c, nn);
    }
    c.m_s = dalloc<double>(c, 2 * nn);
    c.m_t = dalloc<double>(c, 2 * nn);
    c.m_x = dalloc<double>(c, 2 * nn);
    c.c_r = dalloc<double>(c, nn);
    // the tiled mixed pure-Neumann V-cycle (see mg_neu_tiled)
    c.mg_neu_tiled = c.precond == FCFD_PC_MULTIGRID && !c.set_zeroP && c.precision == FCFD_PREC_MIXED &&
                     !c.g.periodic && c.nlev >= 2 && c.mg_tiled;
    if (c.mg_neu_tiled) {
        c.zero_pin = dalloc<uint8_t>(c, nn);
        FCFD_HIP(hipMemset(c.zero_pin, 0, nn));
    }
    if (c.precond == FCFD_PC_MULTIGRID && (c.set_zeroP || c.mg_neu_tiled)) c.c_r2 = dalloc<double>(c, nn);
    c.c_Ap = dalloc<double>(c, nn);
    c.c_x = dalloc<double>(c, nn);
    c.c_rhs = dalloc<double>(c, nn);
    c.ustar = dalloc<double>(c, 2 * nn);
    {
        // the node kernels reduce over up to 2 nn items; the tiled kernels over their tiles
        int ps = blocks_for(2 * nn);
        ps = std::max(ps, tiles_for(g.nnx, g.nny, MOMT_C, MOMT_R));
        ps = std::max(ps, tiles_for(g.nnx, g.nny, CGT_C, CGT_R));
        ps = std::max(ps, tiles_for(g.nnx, g.nny, MGT_UP_C, MGT_UP_R));
        if (c.nlev >= 2) ps = std::max(ps, tiles_for(c.lv[1].g.nnx, c.lv[1].g.nny, MGT_DOWN_CX, MGT_DOWN_CY));
        c.pstride = ps + 1;
    }
    c.partials = dalloc<double>(c, (size_t)MAX_RED * c.pstride);
    c.st = dalloc<DevState>(c, 1);
    FCFD_HIP(hipMemset(c.st, 0, sizeof(DevState)));
    const int go = 1;
    FCFD_HIP(hipMemcpy(&c.st->go, &go, sizeof(int), hipMemcpyHostToDevice));
    FCFD_HIP(hipHostMalloc((void**)&c.h_pinned, 64, hipHostMallocDefault));
    // speculative-step records: written by kernels, read by the host after an event
    FCFD_HIP(hipHostMalloc((void**)&c.h_spec, 8 * sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
    std::memset(c.h_spec, 0, 8 * sizeof(int));
    FCFD_HIP(hipHostGetDevicePointer((void**)&c.d_spec, c.h_spec, 0));
    FCFD_HIP(hipDeviceSynchronize());
}

void destroy_ctx(Ctx& c) {
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    (void)hipDeviceSynchronize();
    for (IterGraph* g : {&c.g_mom, &c.g_cg, &c.g_phi, &c.g_amom, &c.g_aphi}) graph_reset(*g);
    if (c.ev_in) (void)hipEventDestroy(c.ev_in);
    if (c.ev_out) (void)hipEventDestroy(c.ev_out);
    for (auto& e : c.ev_spec) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.ev_in = c.ev_out = nullptr;
    c.stream = nullptr;
    for (void* p : c.allocs) (void)hipFree(p);
    c.allocs.clear();
    if (c.h_pinned) (void)hipHostFree(c.h_pinned);
    c.h_pinned = nullptr;
    if (c.h_spec) (void)hipHostFree(c.h_spec);
    c.h_spec = c.d_spec = nullptr;
}

// ------------------------------------------------------------------------------ speculative steps
// A step is enqueued whole, without a host read: each solve gets a predicted chunk of iterations
// (the device flags make iterations past convergence return at entry), and the last finalize of
// each chunk writes the step's record in host-coherent memory -- the iteration count, or, if the
// solve was still active, that the chunk fell short; it then clears st->go and the solver's flag
// so that every later kernel of this step and of the steps enqueued behind it returns at entry.
// The host reads the record once the step's event has passed and, for a step that fell short,
// resumes the solve where the chunk ended and redoes the rest of the step.
__global__ void k_spec_open(DevState* st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st->go = 1;
}
__global__ void k_spec_resume(DevState* st, int phase) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->go = 1;
    (phase == 0 ? st->b_active : st->c_active) = 1;
}

// chunk length: the larger of the last two iteration counts (0 before any solve: the solver's own
// prediction).  Counts mostly fall or stay level from step to step; a chunk that falls short
// costs a resume and the no-op replay of the step enqueued behind it, one iteration too many
// costs only that iteration's launches (its kernels return at entry).
int spec_chunk(const Ctx& c, const int (&h)[3]) {
    return c.spec_chunk_fixed > 0 ? c.spec_chunk_fixed : std::max(h[0], h[1]);
}
// of such a chunk, the iterations the host is confident of: the smaller of the last two counts
// (-1, all of them, before two solves or with a fixed chunk length)
int spec_sure(const Ctx& c, const int (&h)[3]) {
    if (c.spec_chunk_fixed > 0 || h[0] <= 0 || h[1] <= 0) return -1;
    return std::min(h[0], h[1]);
}
void spec_push(int (&h)[3], int k) {
    h[2] = h[1];
    h[1] = h[0];
    h[0] = k > 0 ? k : 0;
}

struct StepLaunch {
    int slot, mom, p;  // record slot, iterations enqueued per solve
};

// kernels launched while alive take st->go as their gate
struct GateScope {
    Ctx& c;
    explicit GateScope(Ctx& cc) : c(cc) { c.gate = &c.st->go; }
    ~GateScope() {
        c.gate = nullptr;
        c.chunk_sure = -1;
    }
};

// four_step_timestep (standard.py:148-165), enqueued as a speculative step.
// scatter = false: the Dirichlet values are already in place (a rollout scatters them once).
// P is written only by the output pass and U only by the final projection, after both solves,
// so either may alias its input (P == P_n, U == U_n).
StepLaunch four_step_enqueue(Ctx& c, int slot, const double* U_n, const double* P_n, double dt, double rho,
                             double mu, const double* U_d, const double* F_p, const double* F, double* U, double* P,
                             bool scatter) {
    StepLaunch L{slot, spec_chunk(c, c.hist_mom), spec_chunk(c, c.hist_p)};
    // the slot's previous step has been waited for (or did nothing): its record is free
    c.h_spec[4 * slot] = 0;
    int* rec = c.d_spec + 4 * slot;
    if (scatter) scatter_ud(c, U_d);
    {
        GateScope gate(c);
        build_ce(c, U_n, 0.5, c.ce_format, c.ce, c.ce_scale);
        c.chunk_sure = spec_sure(c, c.hist_mom);
        solve_momentum(c, c.ce_format, c.ce, c.ce_scale, U_n, P_n, rho, mu, dt, F, c.m_x, &L.mom, ChunkEnd{rec, 0});
        c.chunk_sure = spec_sure(c, c.hist_p);
        ustar_pressure_solve(c, c.m_x, P_n, F_p, rho, dt, &L.p, ChunkEnd{rec, 1});
        output_and_project(c, P, rho, dt, U);
    }
    FCFD_CHECK_LAUNCH();
    FCFD_HIP(hipEventRecord(c.ev_spec[slot], c.stream));
    return L;
}

// Waits for the step of L.slot and returns its iteration counts in it[0..1].  If a chunk fell
// short, the solve is resumed and the rest of the step redone here (with host reads); steps
// enqueued behind it did nothing and must be enqueued again (the return value is then false).
bool four_step_finish(Ctx& c, const StepLaunch& L, const double* P_n, double dt, double rho, double mu,
                      const double* F_p, double* U, double* P, int* it) {
    FCFD_HIP(hipEventSynchronize(c.ev_spec[L.slot]));
    const volatile int* rec = c.h_spec + 4 * L.slot;
    const int status = rec[0], phase = rec[1], mom_it = rec[2], p_it = rec[3];
    if (status == 1) {
        it[0] = mom_it;
        it[1] = p_it;
    } else {
        if (status != 2) throw Error(FCFD_ERR_STATE, "speculative step found not run");
        k_spec_resume<<<1, 64, 0, c.stream>>>(c.st, phase);
        if (phase == 0) {
            it[0] = resume_momentum(c, c.ce_format, c.ce, c.ce_scale, rho, mu, dt, L.mom);
            it[1] = ustar_pressure_solve(c, c.m_x, P_n, F_p, rho, dt);
        } else {
            it[0] = mom_it;
            it[1] = pressure_resume(c, L.p);
        }
        output_and_project(c, P, rho, dt, U);
        FCFD_CHECK_LAUNCH();
    }
    spec_push(c.hist_mom, it[0]);
    spec_push(c.hist_p, it[1]);
    return status == 1;
}

void four_step(Ctx& c, const double* U_n, const double* P_n, double dt, double rho, double mu, const double* U_d,
               const double* F_p, const double* F, double* U, double* P, int* it, bool scatter = true) {
    k_spec_open<<<1, 64, 0, c.stream>>>(c.st);  // no step is in flight at entry
    const StepLaunch L = four_step_enqueue(c, 0, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, scatter);
    four_step_finish(c, L, P_n, dt, rho, mu, F_p, U, P, it);
}

double* api_ce_buffer(Ctx& c) {
    if (!c.ce64_api) c.ce64_api = dalloc<double>(c, 16 * (size_t)c.g.ne);
    return c.ce64_api;
}

}  // namespace

extern "C" {

const char* fcfd_last_error(void) { return g_last_error.c_str(); }

int fcfd_ctx_create(const fcfd_mesh_desc* mesh, const fcfd_solver_desc* solver, void* stream, fcfd_ctx** out) {
    if (!out) {
        g_last_error = "out is null";
        return FCFD_ERR_ARG;
    }
    *out = nullptr;
    fcfd_ctx* h = new fcfd_ctx();
    int rc = guard([&] {
        build_ctx(h->c, mesh, solver, stream);
        h->scratch = scratch_on_current_device();
    });
    if (rc != FCFD_OK) {
        destroy_ctx(h->c);
        delete h;
        return rc;
    }
    *out = h;
    return FCFD_OK;
}

int fcfd_ctx_destroy(fcfd_ctx* ctx) {
    if (!ctx) return FCFD_OK;
    destroy_ctx(ctx->c);
    delete ctx;
    return FCFD_OK;
}

int fcfd_set_stream(fcfd_ctx* ctx, void* stream) {
    if (!ctx) return FCFD_ERR_ARG;
    ctx->c.user_stream = (hipStream_t)stream;
    return FCFD_OK;
}

int fcfd_ctx_info(const fcfd_ctx* ctx, int64_t* nn, int64_t* ne, int32_t* levels) {
    if (!ctx) return FCFD_ERR_ARG;
    if (nn) *nn = ctx->c.g.nn;
    if (ne) *ne = ctx->c.g.ne;
    if (levels) *levels = ctx->c.nlev;
    return FCFD_OK;
}

int fcfd_level_nodes(const fcfd_ctx* ctx, int32_t level, int64_t* nn, int32_t* nnx, int32_t* nny) {
    if (!ctx || level < 0 || level >= ctx->c.nlev) return FCFD_ERR_ARG;
    const Grid2D& g = ctx->c.lv[level].g;
    if (nn) *nn = g.nn;
    if (nnx) *nnx = g.nnx;
    if (nny) *nny = g.nny;
    return FCFD_OK;
}

int fcfd_mesh_fill(fcfd_ctx* ctx, double* coords, int64_t* nconn) {
    if (!ctx) return FCFD_ERR_ARG;
    return run2d(ctx, "fcfd_mesh_fill", [&] { mesh_fill(ctx->c, coords, nconn); });
}

int fcfd_element_constant(const fcfd_ctx* ctx, int32_t which, double* out16) {
    if (!ctx || !out16) return FCFD_ERR_ARG;
    if (ctx->c.dim != 2 || ctx->c.order != 1) {
        g_last_error = "fcfd_element_constant returns 4 x 4 (quad4) constants; use fcfd_element_matrix";
        return FCFD_ERR_UNSUPPORTED;
    }
    const ElemConst& e = ctx->c.ec;
    const double* src = which == 0 ? e.Me : which == 1 ? e.Le : which == 2 ? e.Gx : which == 3 ? e.Gy
                      : which == 4 ? e.Leoff : nullptr;
    if (!src) {
        g_last_error = "unknown element constant";
        return FCFD_ERR_ARG;
    }
    std::memcpy(out16, src, 16 * sizeof(double));
    return FCFD_OK;
}

int fcfd_element_constant3(const fcfd_ctx* ctx, int32_t which, double* out64) {
    if (!ctx || !out64) return FCFD_ERR_ARG;
    if (ctx->c.dim != 3 || ctx->c.order == 2) {
        g_last_error = "fcfd_element_constant3 needs a hex8 context (hex27: fcfd_element_matrix)";
        return FCFD_ERR_UNSUPPORTED;
    }
    const ElemConst3& e = ctx->c.ec3;
    const double* src = which == 0 ? e.Me : which == 1 ? e.Le : (which >= 2 && which <= 4) ? e.G[which - 2] : nullptr;
    if (!src) {
        g_last_error = "unknown element constant";
        return FCFD_ERR_ARG;
    }
    std::memcpy(out64, src, 64 * sizeof(double));
    return FCFD_OK;
}

int fcfd_element_matrix(const fcfd_ctx* ctx, int32_t which, double* out, int32_t n) {
    if (!ctx || !out) return FCFD_ERR_ARG;
    const Ctx& c = ctx->c;
    const int nen = c.dim == 3 ? (c.order == 2 ? 27 : 8) : (c.order == 2 ? 9 : 4);
    if (n != nen * nen) {
        g_last_error = "element matrix size mismatch: this context has " + std::to_string(nen * nen) + " entries";
        return FCFD_ERR_ARG;
    }
    const double* src = nullptr;
    if (c.dim == 3 && c.order == 2) {
        const ElemConstH27& e = c.eh27[0];
        src = which == 0 ? e.Me : which == 1 ? e.Le : (which >= 2 && which <= 4) ? e.G[which - 2]
            : which == 5 ? e.Leoff : nullptr;
    } else if (c.dim == 3) {
        src = which == 0 ? c.ec3.Me : which == 1 ? c.ec3.Le : (which >= 2 && which <= 4) ? c.ec3.G[which - 2] : nullptr;
    } else if (c.order == 2) {
        src = which == 0 ? c.eq2.Me : which == 1 ? c.eq2.Le : which == 2 ? c.eq2.Gx : which == 3 ? c.eq2.Gy
            : which == 5 ? c.eq2.Leoff : nullptr;
    } else {
        src = which == 0 ? c.ec.Me : which == 1 ? c.ec.Le : which == 2 ? c.ec.Gx : which == 3 ? c.ec.Gy
            : which == 5 ? c.ec.Leoff : nullptr;
    }
    if (!src) {
        g_last_error = "unknown element matrix";
        return FCFD_ERR_ARG;
    }
    std::memcpy(out, src, (size_t)n * sizeof(double));
    return FCFD_OK;
}

int fcfd_mass_matvec(fcfd_ctx* ctx, const double* x, double* y) {
    if (!ctx || !x || !y || x == y) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        Ctx& c = ctx->c;
        if (c.dim == 3 && c.order == 2) {
            mass_matvec_h27(c, x, y);
        } else if (c.dim == 3) {
            mass_matvec3(c, x, y);
        } else if (c.order == 2) {
            mass_matvec_q2(c, x, y);
        } else {
            elem_mult_const(c, 0, c.ec.Me, x, y, 0);
            elem_mult_const(c, 0, c.ec.Me, x + c.g.nn, y + c.g.nn, 0);
        }
    });
}

int fcfd_lumped_mass(fcfd_ctx* ctx, double* M) {
    if (!ctx || !M) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        FCFD_HIP(hipMemcpyAsync(M, ctx->c.Mlump, sizeof(double) * ctx->c.dim * ctx->c.g.nn,
                                hipMemcpyDeviceToDevice, ctx->c.stream));
    });
}

int fcfd_laplacian_diag(fcfd_ctx* ctx, int32_t level, double* out) {
    if (!ctx || !out || level < 0 || level >= ctx->c.nlev) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        const Level& L = ctx->c.lv[level];
        FCFD_HIP(hipMemcpyAsync(out, L.Ld, sizeof(double) * L.g.nn, hipMemcpyDeviceToDevice, ctx->c.stream));
    });
}

int fcfd_conv_elem_calc(fcfd_ctx* ctx, const double* U, double beta, double* Ce) {
    if (!ctx || !U || !Ce) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        if (ctx->c.dim == 3 && ctx->c.order == 2) build_ce_h27_api(ctx->c, U, beta, Ce);
        else if (ctx->c.dim == 3) build_ce3_api(ctx->c, U, beta, Ce);
        else if (ctx->c.order == 2) build_ce_q2_api(ctx->c, U, beta, Ce);
        else build_ce_api(ctx->c, U, beta, Ce);
    });
}

int fcfd_boundary_integral(fcfd_ctx* ctx, const double* U_d, double* F_p) {
    if (!ctx || !F_p) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        if (ctx->c.dim == 3 && ctx->c.order == 2) boundary_integral_h27(ctx->c, U_d, F_p);
        else if (ctx->c.dim == 3) boundary_integral3(ctx->c, U_d, F_p);
        else if (ctx->c.order == 2) boundary_integral_q2(ctx->c, U_d, F_p);
        else boundary_integral(ctx->c, U_d, F_p);
    });
}

int fcfd_elem_mult_const(fcfd_ctx* ctx, int32_t level, const double* A16, const double* x, double* y,
                         int32_t accumulate) {
    if (!ctx || !A16 || !x || !y) return FCFD_ERR_ARG;
    return run2d(ctx, "fcfd_elem_mult_const", [&] { elem_mult_const(ctx->c, level, A16, x, y, accumulate); });
}

int fcfd_elem_mult_stored(fcfd_ctx* ctx, const double* K0_16, const double* A1, double c1, const double* A2,
                          double c2, int32_t ncomp, const double* x, double* y) {
    if (!ctx || !x || !y || ncomp < 1 || ncomp > 3) return FCFD_ERR_ARG;
    return run2d(ctx, "fcfd_elem_mult_stored",
                 [&] { elem_mult_general(ctx->c, K0_16, A1, c1, A2, c2, ncomp, x, y); });
}

int fcfd_vcycle(fcfd_ctx* ctx, const double* r, double* z) {
    if (!ctx || !r || !z) return FCFD_ERR_ARG;
    if (ctx->c.dim == 2 && ctx->c.order == 2) return run(ctx, [&] { vcycle_q2_api(ctx->c, r, z); });
    if (ctx->c.dim == 3 && ctx->c.order == 2) return run(ctx, [&] { vcycle_h27_api(ctx->c, r, z); });
    if (ctx->c.dim == 3) return run(ctx, [&] { vcycle3_api(ctx->c, r, z); });
    return run2d(ctx, "fcfd_vcycle", [&] { vcycle(ctx->c, r, z); });
}

int fcfd_interpolate(fcfd_ctx* ctx, int32_t level, const double* xH, double* xh) {
    if (!ctx || !xH || !xh) return FCFD_ERR_ARG;
    return run2d(ctx, "fcfd_interpolate", [&] { interpolate_level(ctx->c, level, xH, xh); });
}

int fcfd_restrict(fcfd_ctx* ctx, int32_t level, const double* xh, double* xH) {
    if (!ctx || !xh || !xH) return FCFD_ERR_ARG;
    return run2d(ctx, "fcfd_restrict", [&] { restrict_level(ctx->c, level, xh, xH); });
}

int fcfd_jacobi_sweep(fcfd_ctx* ctx, int32_t level, const double* Le_off16, const double* x, const double* rhs,
                      const double* L_diag, const int64_t* gamma_p, int64_t n_gamma_p, int32_t set_zeroP,
                      double* x_out) {
    if (!ctx || !Le_off16 || !x || !rhs || !L_diag || !x_out) return FCFD_ERR_ARG;
    return run2d(ctx, "fcfd_jacobi_sweep", [&] {
        jacobi_sweep(ctx->c, level, Le_off16, x, rhs, L_diag, gamma_p, n_gamma_p, set_zeroP, x_out);
    });
}

int fcfd_pressure_cg(fcfd_ctx* ctx, const double* b, double* x, double tol, int32_t maxiter, int32_t* iters,
                     int32_t* converged) {
    if (!ctx || !b || !x) return FCFD_ERR_ARG;
    if (ctx->c.dim == 3)
        return run(ctx, [&] {
            bool conv = false;
            int it = ctx->c.order == 2 ? pressure_cg_h27(ctx->c, b, x, tol, maxiter > 0 ? maxiter : 1000, &conv)
                                       : pressure_cg3(ctx->c, b, x, tol, maxiter > 0 ? maxiter : 1000, &conv);
            FCFD_HIP(hipStreamSynchronize(ctx->c.stream));
            if (iters) *iters = it;
            if (converged) *converged = conv;
        });
    return run2d(ctx, "fcfd_pressure_cg", [&] {
        bool conv = false;
        int it = pressure_cg(ctx->c, b, x, tol, maxiter > 0 ? maxiter : 1000, &conv);
        FCFD_HIP(hipStreamSynchronize(ctx->c.stream));
        if (iters) *iters = it;
        if (converged) *converged = conv;
    });
}

int fcfd_solve_pressure(fcfd_ctx* ctx, const double* U_star, const double* P_n, const double* F_p, double rho,
                        double dt, double* P, int32_t* iters) {
    if (!ctx || !U_star || !P_n || !P) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        int it = ctx->c.dim == 3 && ctx->c.order == 2 ? solve_pressure_h27(ctx->c, U_star, P_n, F_p, rho, dt, P)
               : ctx->c.dim == 3     ? solve_pressure3(ctx->c, U_star, P_n, F_p, rho, dt, P)
                 : ctx->c.order == 2 ? solve_pressure_q2(ctx->c, U_star, P_n, F_p, rho, dt, P)
                                     : solve_pressure(ctx->c, U_star, P_n, F_p, rho, dt, P);
        FCFD_HIP(hipStreamSynchronize(ctx->c.stream));
        if (iters) *iters = it;
    });
}

int fcfd_solve_momentum(fcfd_ctx* ctx, const double* U_n, const double* P_n, const double* Ce, double rho, double mu,
                        double dt, const double* U_d, const double* F, double* U_hat, int32_t* iters) {
    if (!ctx || !U_n || !P_n || !U_hat) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        Ctx& c = ctx->c;
        scatter_ud(c, U_d);
        int it;
        if (c.dim == 3 && c.order == 2) {
            FCFD_REQUIRE(!Ce, "3D contexts build Ce(U_n) themselves (pass Ce = NULL)");
            it = solve_momentum_h27(c, U_n, P_n, rho, mu, dt, F, U_hat);
        } else if (c.dim == 3) {
            FCFD_REQUIRE(!Ce, "3D contexts build Ce(U_n) themselves (pass Ce = NULL)");
            it = solve_momentum3_from(c, U_n, P_n, rho, mu, dt, F, U_hat);
        } else if (c.order == 2) {
            FCFD_REQUIRE(!Ce, "quad9 contexts build Ce(U_n) themselves (pass Ce = NULL)");
            it = solve_momentum_q2(c, U_n, P_n, rho, mu, dt, F, U_hat);
        } else if (Ce) {
            double* em = api_ce_buffer(c);
            ce_from_api(c, Ce, em);
            it = solve_momentum(c, FCFD_CE_F64, em, nullptr, U_n, P_n, rho, mu, dt, F, U_hat);
        } else {
            build_ce(c, U_n, 0.5, c.ce_format, c.ce, c.ce_scale);
            it = solve_momentum(c, c.ce_format, c.ce, c.ce_scale, U_n, P_n, rho, mu, dt, F, U_hat);
        }
        FCFD_HIP(hipStreamSynchronize(c.stream));
        if (iters) *iters = it;
    });
}

int fcfd_four_step_timestep(fcfd_ctx* ctx, const double* U_n, const double* P_n, double dt, double rho, double mu,
                            const double* U_d, const double* F_p, const double* F, double* U, double* P,
                            int32_t* iters) {
    if (!ctx || !U_n || !P_n || !U || !P) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        int it[2];
        if (ctx->c.dim == 3 && ctx->c.order == 2) four_step_h27(ctx->c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
        else if (ctx->c.dim == 3) four_step3(ctx->c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
        else if (ctx->c.order == 2) four_step_q2(ctx->c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
        else four_step(ctx->c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
        if (iters) {
            iters[0] = it[0];
            iters[1] = it[1];
        }
    });
}

int fcfd_four_step_rollout(fcfd_ctx* ctx, const double* U0, const double* P0, int32_t nsteps, double dt, double rho,
                           double mu, const double* U_d, const double* F_p, const double* F, double* U, double* P,
                           int32_t* iters) {
    if (!ctx || !U0 || !P0 || !U || !P || nsteps < 1) return FCFD_ERR_ARG;
    if (U == U0 || P == P0) return FCFD_ERR_ARG;
    return run(ctx, [&] {
        Ctx& c = ctx->c;
        const size_t nn = (size_t)c.g.nn;
        if (!c.roll_U) {
            c.roll_U = dalloc<double>(c, (size_t)c.dim * nn);
            c.roll_P = dalloc<double>(c, nn);
        }
        // the states alternate between (U, P) and the context's pair so that the last step
        // writes (U, P)
        if (c.dim == 2 && c.order == 1) {
            // one step of lookahead: step k + 1 is enqueued before the record of step k is read,
            // so the device does not wait for the host between steps
            auto out_U = [&](int k) { return (nsteps - 1 - k) % 2 == 0 ? U : c.roll_U; };
            auto out_P = [&](int k) { return (nsteps - 1 - k) % 2 == 0 ? P : c.roll_P; };
            auto in_U = [&](int k) { return k == 0 ? U0 : (const double*)out_U(k - 1); };
            auto in_P = [&](int k) { return k == 0 ? P0 : (const double*)out_P(k - 1); };
            auto enqueue = [&](int k) {
                return four_step_enqueue(c, k & 1, in_U(k), in_P(k), dt, rho, mu, U_d, F_p, F, out_U(k), out_P(k),
                                         k == 0);
            };
            k_spec_open<<<1, 64, 0, c.stream>>>(c.st);  // no step is in flight at entry
            StepLaunch cur = enqueue(0), next{};
            for (int k = 0; k < nsteps; ++k) {
                if (k + 1 < nsteps) next = enqueue(k + 1);
                int it[2];
                const bool full = four_step_finish(c, cur, in_P(k), dt, rho, mu, F_p, out_U(k), out_P(k), it);
                if (!full && k + 1 < nsteps) next = enqueue(k + 1);  // the first one did nothing
                if (iters) {
                    iters[2 * k] = it[0];
                    iters[2 * k + 1] = it[1];
                }
                cur = next;
            }
            return;
        }
        const double *Us = U0, *Ps = P0;
        for (int k = 0; k < nsteps; ++k) {
            const bool to_out = (nsteps - 1 - k) % 2 == 0;
            double* Ud = to_out ? U : c.roll_U;
            double* Pd = to_out ? P : c.roll_P;
            int it[2];
            if (c.dim == 3 && c.order == 2) four_step_h27(c, Us, Ps, dt, rho, mu, U_d, F_p, F, Ud, Pd, it);
            else if (c.dim == 3) four_step3(c, Us, Ps, dt, rho, mu, U_d, F_p, F, Ud, Pd, it);
            else four_step_q2(c, Us, Ps, dt, rho, mu, U_d, F_p, F, Ud, Pd, it);
            if (iters) {
                iters[2 * k] = it[0];
                iters[2 * k + 1] = it[1];
            }
            Us = Ud;
            Ps = Pd;
        }
    });
}

int fcfd_solve_scalar(fcfd_ctx* ctx, const double* U_n, const double* U, const double* phi_n, double dt, double D,
                      const double* phi_d, int32_t supg, double* phi, int32_t* iters) {
    if (!ctx || !U_n || !U || !phi_n || !phi) return FCFD_ERR_ARG;
    if (ctx->c.dim == 2 && ctx->c.order == 2)
        return run(ctx, [&] {
            int it = solve_scalar_q2(ctx->c, U_n, U, phi_n, dt, D, phi_d, supg, phi);
            if (iters) *iters = it;
        });
    return run2d(ctx, "fcfd_solve_scalar", [&] {
        int it = solve_scalar(ctx->c, U_n, U, phi_n, dt, D, phi_d, supg, phi);
        if (iters) *iters = it;
    });
}

int fcfd_passive_timestep(fcfd_ctx* ctx, const double* U_n, const double* P_n, const double* phi_n, double dt,
                          double rho, double mu, double D, const double* U_d, const double* phi_d,
                          const double* F_p, const double* F, int32_t supg, double* U, double* P, double* phi,
                          int32_t* iters) {
    if (!ctx || !U_n || !P_n || !phi_n || !U || !P || !phi) return FCFD_ERR_ARG;
    if (ctx->c.dim == 2 && ctx->c.order == 2)
        return run(ctx, [&] {
            Ctx& c = ctx->c;
            int it[3];
            four_step_q2(c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
            it[2] = solve_scalar_q2(c, U_n, U, phi_n, dt, D, phi_d, supg, phi);
            if (iters)
                for (int k = 0; k < 3; ++k) iters[k] = it[k];
        });
    return run2d(ctx, "fcfd_passive_timestep", [&] {
        Ctx& c = ctx->c;
        int it[3];
        four_step(c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
        it[2] = solve_scalar(c, U_n, U, phi_n, dt, D, phi_d, supg, phi);
        if (iters)
            for (int k = 0; k < 3; ++k) iters[k] = it[k];
    });
}

int fcfd_four_step_timestep_vjp(fcfd_ctx* ctx, const double* U_n, const double* P_n, double dt, double rho,
                                double mu, const double* U_d, const double* F_p, const double* F, const double* gU,
                                const double* gP, double* U, double* P, double* gU_n, double* gP_n, int32_t* iters) {
    if (!ctx || !U_n || !P_n || !U || !P || !gU_n || !gP_n) return FCFD_ERR_ARG;
    if (U == U_n || P == P_n || gU_n == U_n || gP_n == P_n) return FCFD_ERR_ARG;
    if (ctx->c.dim == 2 && ctx->c.order == 2)
        return run(ctx, [&] {
            Ctx& c = ctx->c;
            int it[4];
            four_step_q2(c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
            ns_vjp_q2(c, U_n, P_n, dt, rho, mu, gU, gP, nullptr, gU_n, gP_n, it + 2);
            if (iters)
                for (int k = 0; k < 4; ++k) iters[k] = it[k];
        });
    return run2d(ctx, "fcfd_four_step_timestep_vjp", [&] {
        Ctx& c = ctx->c;
        int it[4];
        four_step(c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
        ns_vjp(c, U_n, P_n, dt, rho, mu, gU, gP, nullptr, gU_n, gP_n, it + 2);
        if (iters)
            for (int k = 0; k < 4; ++k) iters[k] = it[k];
    });
}

int fcfd_passive_timestep_vjp(fcfd_ctx* ctx, const double* U_n, const double* P_n, const double* phi_n, double dt,
                              double rho, double mu, double D, const double* U_d, const double* phi_d,
                              const double* F_p, const double* F, int32_t supg, const double* gU, const double* gP,
                              const double* gphi, double* U, double* P, double* phi, double* gU_n, double* gP_n,
                              double* gphi_n, int32_t* iters) {
    if (!ctx || !U_n || !P_n || !phi_n || !U || !P || !phi || !gU_n || !gP_n || !gphi_n) return FCFD_ERR_ARG;
    if (U == U_n || P == P_n || phi == phi_n || gU_n == U_n || gP_n == P_n || gphi_n == phi_n) return FCFD_ERR_ARG;
    if (ctx->c.dim == 2 && ctx->c.order == 2) {
        return run(ctx, [&] {
            Ctx& c = ctx->c;
            int it[6];
            four_step_q2(c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
            it[2] = solve_scalar_q2(c, U_n, U, phi_n, dt, D, phi_d, supg, phi);
            passive_vjp_q2(c, U_n, P_n, phi_n, U, phi, dt, rho, mu, D, supg, gU, gP, gphi, gU_n, gP_n, gphi_n, it + 3);
            if (iters)
                for (int k = 0; k < 6; ++k) iters[k] = it[k];
        });
    }
    return run2d(ctx, "fcfd_passive_timestep_vjp", [&] {
        Ctx& c = ctx->c;
        int it[6];
        four_step(c, U_n, P_n, dt, rho, mu, U_d, F_p, F, U, P, it);
        it[2] = solve_scalar(c, U_n, U, phi_n, dt, D, phi_d, supg, phi);
        passive_vjp(c, U_n, P_n, phi_n, U, phi, dt, rho, mu, D, supg, gU, gP, gphi, gU_n, gP_n, gphi_n, it + 3);
        if (iters)
            for (int k = 0; k < 6; ++k) iters[k] = it[k];
    });
}

int fcfd_supg_elem_calc(fcfd_ctx* ctx, const double* U_n, const double* Ce, double D, double dt, double r, double* Ct,
                        double* tau) {
    if (!ctx || !U_n) return FCFD_ERR_ARG;
    return run2d(ctx, "fcfd_supg_elem_calc", [&] { supg_api(ctx->c, U_n, Ce, D, dt, r, Ct, tau); });
}

int fcfd_dot(fcfd_ctx* ctx, const double* x, const double* y, int64_t n, double* out) {
    if (!ctx || !x || !y || !out || n < 0) return FCFD_ERR_ARG;
    return run(ctx, [&] { *out = dot(ctx->c, x, y, (int)n); });
}

int fcfd_dot_noalloc(fcfd_ctx* ctx, const double* x, const double* y, int64_t n, double* out) {
    if (!ctx || !x || !y || !out || n < 0) return FCFD_ERR_ARG;
    if (n > FCFD_DOT_NOALLOC_MAX_N) {
        g_last_error = "fcfd_dot_noalloc: n = " + std::to_string(n) + " exceeds FCFD_DOT_NOALLOC_MAX_N = " +
                       std::to_string(FCFD_DOT_NOALLOC_MAX_N) + " (the size of its fixed partials array)";
        return FCFD_ERR_ARG;
    }
    return run(ctx, [&] { dot_to_device(ctx->c, x, y, (int)n, ctx->scratch, out); });
}

int fcfd_axpby(fcfd_ctx* ctx, double a, const double* x, double b, double* y, int64_t n) {
    if (!ctx || !x || !y || n < 0) return FCFD_ERR_ARG;
    return run(ctx, [&] { axpby(ctx->c, a, x, b, y, (int)n); });
}

int fcfd_bench_momentum_apply(fcfd_ctx* ctx, int32_t reps, float* ms_per_apply) {
    if (!ctx || reps <= 0 || !ms_per_apply) return FCFD_ERR_ARG;
    return run2d(ctx, "fcfd_bench_momentum_apply", [&] { *ms_per_apply = bench_momentum_apply(ctx->c, reps); });
}

}  // extern "C"
