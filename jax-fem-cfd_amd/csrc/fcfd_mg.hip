// Pressure solve: geometric multigrid V-cycle (jax_fem_cfd/core/multigrid.py:231-330) as a
// preconditioner of CG (jax_fem_cfd/utils/jax_iterative_solvers.py:87-120), and
// solve_pressure (jax_fem_cfd/solvers/standard.py:85-125).
//
// V-cycle kernels per level (pinned form / pure-Neumann form):
//   down   [DownX (Neumann)], DownRes, Restrict
//   coarse one single-workgroup kernel: CG to tol (pinned) or one Jacobi (Neumann)
//   up     [Prolong (Neumann)], Jacobi (prolongation formed on the fly in the pinned form)
// The pure-Neumann mean projections are applied lazily: every producing kernel reduces the sum
// of what it writes (fused, deterministic) and consumers subtract the mean when they read.
// Level vectors are templated on T: double (plain FP64 form) or float (mixed form).
#include "fcfd_internal.hpp"
#include "fcfd_ops.hpp"
#include "fcfd_cg.hpp"
#include "fcfd_mgnodes.hpp"
#include "fcfd_mgtile.hpp"

namespace fcfd {


template <class T> struct LvBufs;
template <> struct LvBufs<double> {
    static double* r(Level& L) { return L.r; }
    static double* x(Level& L) { return L.x; }
    static double* y(Level& L) { return L.y; }
    static double* f(Level& L) { return L.f; }
    static const double* Ld(Level& L) { return L.Ld; }
    static double* cr(Ctx& c) { return c.cc_r; }
    static double* cp(Ctx& c) { return c.cc_p; }
    static double* cAp(Ctx& c) { return c.cc_Ap; }
};
template <> struct LvBufs<float> {
    static float* r(Level& L) { return L.r32; }
    static float* x(Level& L) { return L.x32; }
    static float* y(Level& L) { return L.y32; }
    static float* f(Level& L) { return L.f32; }
    static const float* Ld(Level& L) { return L.Ld32; }
    static float* cr(Ctx& c) { return c.cc_r32; }
    static float* cp(Ctx& c) { return c.cc_p32; }
    static float* cAp(Ctx& c) { return c.cc_Ap32; }
};

// ------------------------------------------------------------------------------ down leg
// rH = rh - L xh   (Neumann: multigrid.py:263-264; pinned: :246-251)
template <class T, class TR>
struct DownRes {
    Grid2D g; const double* Le; const TR* r; const T* Ld; const uint8_t* pin; T* x; T* f;
    const DevState* st; int l; int neumann;
    __device__ void operator()(int n, double (&acc)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        T X[3][3];
        T fv;
        if (neumann) {
            const T mx = (T)st->mg[l][MG_X], mr = (T)st->mg[l][MG_R];
            load3x3(s, [&](int k) { return x[k] - mx; }, X);
            fv = ((T)r[n] - mr) - ebe_const(s, Le, X);
        } else {
            if (pin)
                load3x3(s, [&](int k) { const T v = (T(2.0 / 3.0) * (T)r[k]) / Ld[k]; return pin[k] ? T(0) : v; }, X);
            else
                load3x3(s, [&](int k) { return (T(2.0 / 3.0) * (T)r[k]) / Ld[k]; }, X);
            x[n] = X[1][1];
            const bool pn = pin && pin[n];
            const T rh = pn ? T(0) : (T)r[n];
            const T lx = pn ? T(0) : ebe_const(s, Le, X);
            fv = rh - lx;
        }
        f[n] = fv;
        acc[0] += (double)fv;
    }
};

// LDS-staged form (non-periodic grids): the smoothed x (pinned: 2/3 r / L_diag, 0 at the pins;
// Neumann: x - mean x) is staged once per node of the block window instead of being formed at
// each of the 9 stencil points; same values bit for bit.
template <class T, class TR>
struct DownResS {
    static constexpr int NV = 1;
    static constexpr int NPT = 1;
    Grid2D g; const double* Le; const TR* r; const T* Ld; const uint8_t* pin; T* x; T* f;
    const DevState* st; int l; int neumann;
    __device__ __forceinline__ int nnx() const { return g.nnx; }
    __device__ __forceinline__ T xval(int k) const {
        if (neumann) return x[k] - (T)st->mg[l][MG_X];
        const T v = (T(2.0 / 3.0) * (T)r[k]) / Ld[k];
        return (pin && pin[k]) ? T(0) : v;
    }
    __device__ __forceinline__ void stage(int id, double (&v)[1]) const { v[0] = (double)xval(id); }
    __device__ void operator()(int n, const StageView<1, NPT>& sv, double (&acc)[1]) const {
        NodeNbr s;
        node_elems(g, n, s);
        T X[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) X[a][b] = (T)sv.at(0, n, a - 1, b - 1);
        T fv;
        if (neumann) {
            const T mr = (T)st->mg[l][MG_R];
            fv = ((T)r[n] - mr) - ebe_const(s, Le, X);
        } else {
            x[n] = X[1][1];
            const bool pn = pin && pin[n];
            const T rh = pn ? T(0) : (T)r[n];
            const T lx = pn ? T(0) : ebe_const(s, Le, X);
            fv = rh - lx;
        }
        f[n] = fv;
        acc[0] += (double)fv;
    }
};

// ------------------------------------------------------------------------------ fused down leg
// Pinned form, non-periodic grids: DownRes and Restrict of one level in one kernel.  A block owns
// DF_TC coarse nodes of one coarse row J; it stages x = 2/3 r / L_diag (0 at the pins and outside
// the grid) on the fine window rows 2J-2 .. 2J+2, writes the x of the fine nodes it owns (rows
// 2J, 2J+1), forms f = r - L x on rows 2J-1 .. 2J+1 in LDS (recomputing the halo of its
// neighbours) and restricts f into its coarse nodes.  f never goes to memory: in the pinned form
// only the restriction reads it.  Values and operation order are those of DownRes + Restrict.
constexpr int DF_TC = 64;
// DR coarse rows per block: the block stages fine rows 2 J0 - 2 .. 2 J0 + 2 DR (2 DR + 3 rows for
// 2 DR owned ones), so larger DR reads fewer halo rows twice.  Every staging load of a thread is
// issued before the first LDS store.
template <class T, class TR, int DR>
__global__ __launch_bounds__(256) void k_down_fused(Grid2D gf, Grid2D gc, const double* __restrict__ Le,
                                                    const TR* __restrict__ r, const T* __restrict__ Ld,
                                                    const uint8_t* __restrict__ pin, T* __restrict__ x,
                                                    T* __restrict__ rc, const int* __restrict__ active) {
    if (active && *active == 0) return;
    constexpr int XW = 2 * DF_TC + 3, FW = 2 * DF_TC + 1;
    constexpr int XR = 2 * DR + 3, FR = 2 * DR + 1;
    __shared__ T xs[XR][XW];
    __shared__ T fs[FR][FW];
    // XCD-aware: each XCD takes a contiguous range of coarse rows, so the fine halo rows that
    // neighbouring blocks both stage are read through the same L2
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int J0 = (lb / nbx) * DR, I0 = (lb - (lb / nbx) * nbx) * DF_TC;
    const int fi0 = 2 * I0 - 2, fj0 = 2 * J0 - 2;
    constexpr int NIT = (XR * XW + 255) / 256;
    T v[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int t = threadIdx.x + it * 256;
        const int a = t / XW, b = t - a * XW;
        const int fi = fi0 + b, fj = fj0 + a;
        const bool in = t < XR * XW && fi >= 0 && fi < gf.nnx && fj >= 0 && fj < gf.nny;
        const int k = in ? fj * gf.nnx + fi : 0;
        const T xv = (T(2.0 / 3.0) * (T)r[k]) / Ld[k];
        v[it] = (!in || (pin && pin[k])) ? T(0) : xv;
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int t = threadIdx.x + it * 256;
        if (t < XR * XW) {
            const int a = t / XW, b = t - a * XW;
            xs[a][b] = v[it];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * DR * 2 * DF_TC; t += 256) {
        const int a = t / (2 * DF_TC), b = t - a * (2 * DF_TC);
        const int fi = 2 * I0 + b, fj = 2 * J0 + a;
        if (fi < gf.nnx && fj < gf.nny) x[fj * gf.nnx + fi] = xs[a + 2][b + 2];
    }
    for (int t = threadIdx.x; t < FR * FW; t += 256) {
        const int a = t / FW, b = t - a * FW;
        const int fi = 2 * I0 - 1 + b, fj = 2 * J0 - 1 + a;
        T fv = T(0);
        if (fi >= 0 && fi < gf.nnx && fj >= 0 && fj < gf.nny) {
            const int n = fj * gf.nnx + fi;
            NodeNbr s;
            node_elems(gf, n, s);
            T X[3][3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) X[dy][dx] = xs[a + dy][b + dx];
            const bool pn = pin && pin[n];
            const T rh = pn ? T(0) : (T)r[n];
            const T lx = pn ? T(0) : ebe_const(s, Le, X);
            fv = rh - lx;
        }
        fs[a][b] = fv;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < DR * DF_TC; t += 256) {
        const int jr = t / DF_TC, ci = t - jr * DF_TC;
        const int I = I0 + ci, J = J0 + jr;
        if (I < gc.nnx && J < gc.nny) {
            const int cc = 2 * ci + 1, rr = 2 * jr + 1;
            auto val = [&](int dy, int dx) -> T { return fs[rr + dy][cc + dx]; };
            const T E = ((val(1, 0) + val(-1, 0)) + val(0, 1)) + val(0, -1);
            const T D = ((val(1, 1) + val(1, -1)) + val(-1, 1)) + val(-1, -1);
            rc[J * gc.nnx + I] = (val(0, 0) + T(0.5) * E) + T(0.25) * D;
        }
    }
}

// Two down legs in one kernel (small pinned levels l -> m = l + 1 -> c = l + 2, where each leg
// alone is a few microseconds of launch and memory latency).  A block owns D2_TR x D2_TC nodes of
// level c and recomputes what they depend on: the level-m residual f_m on the (2 D2_TR + 1) x
// (2 D2_TC + 1) m-nodes around them, hence x_m and r_m on a one-wider m-window, r_m being the
// restriction of f_l on the (2 (2 D2_TR + 3) + 1)-row l-window, hence x_l on a one-wider l-window.
// It stores x_l, r_m and x_m of the nodes it owns (the ones the single-level kernel's block
// would own) and r_c.  Every value is formed with the single-level kernel's operands and order,
// so all four outputs are bitwise those of two k_down_fused launches.
constexpr int D2_TR = 2, D2_TC = 16;
template <class T, class TR>
__global__ __launch_bounds__(256) void k_down2(Grid2D gf, Grid2D gm, Grid2D gc, const double* __restrict__ Lef,
                                               const double* __restrict__ Lem, const TR* __restrict__ rf,
                                               const T* __restrict__ Ldf, const uint8_t* __restrict__ pinf,
                                               T* __restrict__ xf, const T* __restrict__ Ldm,
                                               const uint8_t* __restrict__ pinm, T* __restrict__ rm,
                                               T* __restrict__ xm, T* __restrict__ rc, const int* __restrict__ active) {
    if (active && *active == 0) return;
    constexpr int XRm = 2 * D2_TR + 3, XWm = 2 * D2_TC + 3;  // m-window of x_m, r_m
    constexpr int FRm = 2 * D2_TR + 1, FWm = 2 * D2_TC + 1;  // m-window of f_m
    constexpr int FRl = 2 * XRm + 1, FWl = 2 * XWm + 1;      // l-window of f_l
    constexpr int XRl = FRl + 2, XWl = FWl + 2;              // l-window of x_l
    __shared__ T xl[XRl][XWl];
    __shared__ T fl[FRl][FWl];
    __shared__ T xms[XRm][XWm];
    __shared__ T rms[XRm][XWm];
    __shared__ uint8_t pms[XRm][XWm];
    __shared__ T fm[FRm][FWm];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int J0 = (lb / nbx) * D2_TR, I0 = (lb - (lb / nbx) * nbx) * D2_TC;
    const int mj0 = 2 * J0 - 2, mi0 = 2 * I0 - 2;    // origin of the x_m window
    const int lj0 = 2 * mj0 - 2, li0 = 2 * mi0 - 2;  // origin of the x_l window (f_l: one in)
    // x_l = (2/3 r_l) / L_diag, 0 at the pins and outside the grid
    constexpr int NIT = (XRl * XWl + 255) / 256;
    T v[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int t = threadIdx.x + it * 256;
        const int a = t / XWl, b = t - a * XWl;
        const int fi = li0 + b, fj = lj0 + a;
        const bool in = t < XRl * XWl && fi >= 0 && fi < gf.nnx && fj >= 0 && fj < gf.nny;
        const int k = in ? fj * gf.nnx + fi : 0;
        const T xv = (T(2.0 / 3.0) * (T)rf[k]) / Ldf[k];
        v[it] = (!in || (pinf && pinf[k])) ? T(0) : xv;
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int t = threadIdx.x + it * 256;
        if (t < XRl * XWl) {
            const int a = t / XWl, b = t - a * XWl;
            xl[a][b] = v[it];
        }
    }
    __syncthreads();
    // owned x_l: l-rows 4 J0 .. 4 J0 + 4 D2_TR - 1 (window rows 6 ..), same for columns
    for (int t = threadIdx.x; t < 4 * D2_TR * 4 * D2_TC; t += 256) {
        const int a = t / (4 * D2_TC), b = t - a * (4 * D2_TC);
        const int fi = 4 * I0 + b, fj = 4 * J0 + a;
        if (fi < gf.nnx && fj < gf.nny) xf[fj * gf.nnx + fi] = xl[a + 6][b + 6];
    }
    // f_l = r_l - L x_l (pinned rows 0), 0 outside the grid
    for (int t = threadIdx.x; t < FRl * FWl; t += 256) {
        const int a = t / FWl, b = t - a * FWl;
        const int fi = li0 + 1 + b, fj = lj0 + 1 + a;
        T fv = T(0);
        if (fi >= 0 && fi < gf.nnx && fj >= 0 && fj < gf.nny) {
            const int n = fj * gf.nnx + fi;
            NodeNbr s;
            node_elems(gf, n, s);
            T X[3][3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) X[dy][dx] = xl[a + dy][b + dx];
            const bool pn = pinf && pinf[n];
            const T rh = pn ? T(0) : (T)rf[n];
            const T lx = pn ? T(0) : ebe_const(s, Lef, X);
            fv = rh - lx;
        }
        fl[a][b] = fv;
    }
    __syncthreads();
    // r_m = R f_l on the x_m window, x_m from it; owned r_m, x_m stored
    for (int t = threadIdx.x; t < XRm * XWm; t += 256) {
        const int a = t / XWm, b = t - a * XWm;
        const int mi = mi0 + b, mj = mj0 + a;
        const bool in = mi >= 0 && mi < gm.nnx && mj >= 0 && mj < gm.nny;
        T rv = T(0), xv = T(0);
        bool pk = false;
        if (in) {
            const int rr = 2 * a + 1, cc = 2 * b + 1;
            auto val = [&](int dy, int dx) -> T { return fl[rr + dy][cc + dx]; };
            const T E = ((val(1, 0) + val(-1, 0)) + val(0, 1)) + val(0, -1);
            const T D = ((val(1, 1) + val(1, -1)) + val(-1, 1)) + val(-1, -1);
            rv = (val(0, 0) + T(0.5) * E) + T(0.25) * D;
            const int k = mj * gm.nnx + mi;
            pk = pinm && pinm[k];
            const T x = (T(2.0 / 3.0) * rv) / Ldm[k];
            xv = pk ? T(0) : x;
            if (a >= 2 && a < 2 + 2 * D2_TR && b >= 2 && b < 2 + 2 * D2_TC) {
                rm[k] = rv;
                xm[k] = xv;
            }
        }
        rms[a][b] = rv;
        xms[a][b] = xv;
        pms[a][b] = pk;
    }
    __syncthreads();
    // f_m = r_m - L x_m (pinned rows 0), 0 outside the grid
    for (int t = threadIdx.x; t < FRm * FWm; t += 256) {
        const int a = t / FWm, b = t - a * FWm;
        const int mi = mi0 + 1 + b, mj = mj0 + 1 + a;
        T fv = T(0);
        if (mi >= 0 && mi < gm.nnx && mj >= 0 && mj < gm.nny) {
            NodeNbr s;
            node_elems(gm, mj * gm.nnx + mi, s);
            T X[3][3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) X[dy][dx] = xms[a + dy][b + dx];
            const bool pn = pms[a + 1][b + 1];
            const T rh = pn ? T(0) : rms[a + 1][b + 1];
            const T lx = pn ? T(0) : ebe_const(s, Lem, X);
            fv = rh - lx;
        }
        fm[a][b] = fv;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < D2_TR * D2_TC; t += 256) {
        const int jr = t / D2_TC, ci = t - jr * D2_TC;
        const int I = I0 + ci, J = J0 + jr;
        if (I < gc.nnx && J < gc.nny) {
            const int cc = 2 * ci + 1, rr = 2 * jr + 1;
            auto val = [&](int dy, int dx) -> T { return fm[rr + dy][cc + dx]; };
            const T E = ((val(1, 0) + val(-1, 0)) + val(0, 1)) + val(0, -1);
            const T D = ((val(1, 1) + val(1, -1)) + val(-1, 1)) + val(-1, -1);
            rc[J * gc.nnx + I] = (val(0, 0) + T(0.5) * E) + T(0.25) * D;
        }
    }
}

template <class T, class TR>
static void launch_down_fused(Ctx& c, const Grid2D& gf, const Grid2D& gc, const double* Le, const TR* r, const T* Ld,
                              const uint8_t* pin, T* x, T* rc, const int* active) {
    const int nbx = (gc.nnx + DF_TC - 1) / DF_TC;
    if (gc.nny >= 200) {
        constexpr int DR = 4;
        k_down_fused<T, TR, DR><<<dim3(nbx, (gc.nny + DR - 1) / DR), 256, 0, c.stream>>>(gf, gc, Le, r, Ld, pin, x,
                                                                                       rc, active);
    } else {
        k_down_fused<T, TR, 1><<<dim3(nbx, gc.nny), 256, 0, c.stream>>>(gf, gc, Le, r, Ld, pin, x, rc, active);
    }
}

// ------------------------------------------------------------------------------ up leg
// Pinned-form up-leg value at node n: x + I y_coarse on the stencil (0 at missing and pinned
// nodes), then one damped Jacobi, 0 at the pins -- Jacobi's pinned path with its operands and
// operation order.  yv reads the coarse y (see interp3x3_v).
template <class T, class TR, class YV>
__device__ __forceinline__ T jacobi_pinned(const Grid2D& gf, const Grid2D& gc, const double* Leoff, const T* Ld,
                                           const uint8_t* pin, const TR* r, const T* x, int n, YV yv) {
    NodeNbr s;
    node_nbr(gf, n, s);
    T Y[3][3];
    interp3x3_v(gc, yv, T(0), s, Y);
    T X[3][3];
    load3x3(s, [&](int k) { return x[k]; }, X);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const bool keep = s.ok[a][b] && !(pin && pin[s.nb[a][b]]);
            X[a][b] = keep ? X[a][b] + Y[a][b] : T(0);
        }
    const T rh = (T)r[n];
    const T off = ebe_const(s, Leoff, X);
    T v = (T(OMEGA) * (rh - off)) / Ld[n] + T(1.0 - OMEGA) * X[1][1];
    if (pin && pin[n]) v = T(0);
    return v;
}

// xh <- xh + I xH; one damped Jacobi (multigrid.py:214-228, :306-326)
template <class T, class TR>
struct Jacobi {
    Grid2D gf, gc; const double* Leoff; const T* Ld; const uint8_t* pin; const TR* r; const T* x; const T* p;
    const T* yc; T* y; const DevState* st; int l; int neumann;
    __device__ void operator()(int n, double (&acc)[1]) const {
        NodeNbr s;
        node_nbr(gf, n, s);
        T X[3][3];
        T rh;
        if (neumann) {
            const T mx = (T)st->mg[l][MG_X], mp = (T)st->mg[l][MG_P];
            load3x3(s, [&](int k) { return (x[k] - mx) + (p[k] - mp); }, X);
            const T oup = l == 0 ? T(0) : (T)st->mg[l][MG_R];
            rh = (T)r[n] - oup;
        } else {
            T Y[3][3];
            interp3x3(gc, yc, T(0), s, Y);
            load3x3(s, [&](int k) { return x[k]; }, X);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const bool keep = s.ok[a][b] && !(pin && pin[s.nb[a][b]]);
                    X[a][b] = keep ? X[a][b] + Y[a][b] : T(0);
                }
            rh = (T)r[n];
        }
        const T off = ebe_const(s, Leoff, X);
        T v = (T(OMEGA) * (rh - off)) / Ld[n] + T(1.0 - OMEGA) * X[1][1];
        if (!neumann && pin && pin[n]) v = T(0);
        y[n] = v;
        acc[0] += (double)v;
    }
};

// Two up legs in one kernel (small pinned levels c = l + 2 -> m = l + 1 -> l): a block owns a
// U2_LR x U2_LC tile of level-l nodes; it first forms the level-m Jacobi on the m-window that the
// tile's prolongation patches read (rows (j0 - 1) >> 1 .. + U2_LR / 2 + 2, likewise columns)
// into LDS, then the level-l Jacobi of its nodes with the prolongation read from that window.
// Level m's y is not stored (only level l's prolongation reads it).  The values are those of
// two Jacobi launches (jacobi_pinned, same operands and order).
constexpr int U2_LR = 4, U2_LC = 64;
template <class T, class TR>
__global__ __launch_bounds__(256) void k_up2(Grid2D gf, Grid2D gm, Grid2D gc, const double* __restrict__ Leoff_f,
                                             const double* __restrict__ Leoff_m, const T* __restrict__ Ldf,
                                             const uint8_t* __restrict__ pinf, const TR* __restrict__ rf,
                                             const T* __restrict__ xf, const T* __restrict__ Ldm,
                                             const uint8_t* __restrict__ pinm, const T* __restrict__ rm,
                                             const T* __restrict__ xm, const T* __restrict__ yc, T* __restrict__ yf,
                                             const int* __restrict__ active) {
    if (active && *active == 0) return;
    constexpr int MR = U2_LR / 2 + 3, MC = U2_LC / 2 + 3;
    __shared__ T ym[MR][MC];
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    const int j0 = (lb / nbx) * U2_LR, i0 = (lb - (lb / nbx) * nbx) * U2_LC;
    const int mj0 = (j0 - 1) >> 1, mi0 = (i0 - 1) >> 1;
    for (int t = threadIdx.x; t < MR * MC; t += 256) {
        const int a = t / MC, b = t - a * MC;
        const int mi = mi0 + b, mj = mj0 + a;
        T v = T(0);
        if (mi >= 0 && mi < gm.nnx && mj >= 0 && mj < gm.nny)
            v = jacobi_pinned(gm, gc, Leoff_m, Ldm, pinm, rm, xm, mj * gm.nnx + mi,
                              [&](int cy, int cx) { return yc[cy * gc.nnx + cx]; });
        ym[a][b] = v;
    }
    __syncthreads();
    const int j = j0 + threadIdx.x / U2_LC, i = i0 + threadIdx.x % U2_LC;
    if (i < gf.nnx && j < gf.nny)
        yf[j * gf.nnx + i] = jacobi_pinned(gf, gm, Leoff_f, Ldf, pinf, rf, xf, j * gf.nnx + i,
                                           [&](int cy, int cx) { return ym[cy - mj0][cx - mi0]; });
}

// LDS-staged form of the pinned Jacobi (non-periodic grids): x and the pin flags are staged per
// block window; the operands and operation order are those of Jacobi, so every value is bitwise
// the same.
template <class T, class TR>
struct JacobiS {
    static constexpr int NV = 2;
    static constexpr int NPT = 1;
    Grid2D gf, gc; const double* Leoff; const T* Ld; const uint8_t* pin; const TR* r; const T* x; const T* yc; T* y;
    __device__ __forceinline__ int nnx() const { return gf.nnx; }
    __device__ __forceinline__ void stage(int id, double (&v)[2]) const {
        v[0] = (double)x[id];
        v[1] = (pin && pin[id]) ? 1.0 : 0.0;
    }
    __device__ __forceinline__ T value(int n, const StageView<2, 1>& sv) const {
        NodeNbr s;
        node_nbr(gf, n, s);
        T Y[3][3];
        interp3x3(gc, yc, T(0), s, Y);
        T X[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const bool keep = s.ok[a][b] && sv.at(1, n, a - 1, b - 1) == 0.0;
                X[a][b] = keep ? (T)sv.at(0, n, a - 1, b - 1) + Y[a][b] : T(0);
            }
        const T off = ebe_const(s, Leoff, X);
        T v = (T(OMEGA) * ((T)r[n] - off)) / Ld[n] + T(1.0 - OMEGA) * X[1][1];
        if (sv.at(1, n, 0, 0) != 0.0) v = T(0);
        return v;
    }
    __device__ void operator()(int n, const StageView<2, 1>& sv, double (&)[1]) const { y[n] = value(n, sv); }
};

// The PCG's z-step fused into the last kernel of the pinned V-cycle (level 0, staged): the
// preconditioned z = y is stored as FP64 into the CG's z buffer and the block partials of r.z
// feed the CG's finalize, as the separate pass would (CgK3: r.z and, flexible form, r.z_old;
// CgGamma0 before the first iteration: r.z, with p_old cleared).  The same values, summed per
// block in the same tree over the same node blocks, so gamma and beta are bitwise those of the
// separate pass.
template <class T, int NR>
struct JacobiSZ {
    static constexpr int NV = 2;
    static constexpr int NPT = 1;
    JacobiS<T, double> j;
    double* znew; const double* zold; int flex; double* pzero;
    __device__ __forceinline__ int nnx() const { return j.nnx(); }
    __device__ __forceinline__ void stage(int id, double (&v)[2]) const { j.stage(id, v); }
    __device__ void operator()(int n, const StageView<2, 1>& sv, double (&acc)[NR]) const {
        const double zk = (double)j.value(n, sv);
        const double rk = j.r[n];
        znew[n] = zk;
        if (pzero) pzero[n] = 0.0;
        acc[0] += rk * zk;
        if constexpr (NR == 2) {
            if (flex) acc[1] += rk * zold[n];
        }
    }
};

// CG tail handed to the V-cycle (see JacobiSZ); first: the pass before the first iteration.
// With cg_p set (tiled V-cycle, see cg_fusable) the iteration's x / r update also rides on the
// first down leg: r_new = r0 - alpha Ap goes to cg_rout, x += alpha p, and the finalize of the
// z-step takes r_new . r_new as well.
struct CgTail {
    double* znew; const double* zold; double* pzero; int flex; int first;
    const double* cg_p = nullptr; const double* cg_Ap = nullptr; double* cg_x = nullptr; double* cg_rout = nullptr;
    ChunkEnd ce{};  // the iteration ends a speculative chunk: its last finalize records it
    // z is stored in FP32 (the tiled FP32 V-cycle's z-step; the z buffers then hold floats): set
    // only where the tiled K1 reads it (z_f32), FP64 everywhere else
    int z32 = 0;
};

template <class X>
struct type_tag {
    using type = X;
};

// ------------------------------------------------------------------------------ coarse level
// One workgroup.  Pinned form: CG (M = identity) to tol with maxiter 1000 on L with the pinned
// rows zeroed (multigrid.py:278-287).  Neumann form: one Jacobi from zero (:289-301).

// Pinned, coarse grid of at most NPT * blockDim nodes: search direction in LDS, x / r / Ap in
// registers (node n = tid + i * blockDim, i < NPT).
template <class T, class TR, int NPT>
__global__ __launch_bounds__(1024) void k_coarse_cg_lds(Grid2D g, const ElemConst* __restrict__ ec,
                                                        const uint8_t* pin, const TR* r, T* x_out, DevState* st,
                                                        double tol, int maxit, const int* active) {
    if (active && *active == 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* p = reinterpret_cast<T*>(smem_raw);
    const int nn = g.nn, bs = blockDim.x;
    T xr[NPT], rv[NPT], ap[NPT];
    double g2[1] = {0.0};
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int n = threadIdx.x + i * bs;
        xr[i] = T(0); rv[i] = T(0); ap[i] = T(0);
        if (n < nn) {
            const T b = (pin && pin[n]) ? T(0) : (T)r[n];
            rv[i] = b;
            p[n] = b;
            g2[0] += (double)b * (double)b;
        }
    }
    block_sum<1>(g2);  // barrier: p complete
    double gamma = g2[0];
    const double atol2 = fmax(tol * tol * gamma, 0.0);
    int k = 0;
    while (gamma > atol2 && k < maxit) {
        double pap[1] = {0.0};
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            const int n = threadIdx.x + i * bs;
            if (n < nn) {
                NodeNbr s;
                node_nbr(g, n, s);
                T X[3][3];
                load3x3(s, [&](int m) { return p[m]; }, X);
                const T a = (pin && pin[n]) ? T(0) : ebe_const(s, ec->Le, X);
                ap[i] = a;
                pap[0] += (double)X[1][1] * (double)a;
            }
        }
        block_sum<1>(pap);  // barrier: every read of p by the matvec is done
        const double alpha = gamma / pap[0];
        double rr[1] = {0.0};
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            const int n = threadIdx.x + i * bs;
            if (n < nn) {
                xr[i] = xr[i] + (T)alpha * p[n];
                rv[i] = rv[i] - (T)alpha * ap[i];
                rr[0] += (double)rv[i] * (double)rv[i];
            }
        }
        block_sum<1>(rr);
        const double beta = rr[0] / gamma;
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            const int n = threadIdx.x + i * bs;
            if (n < nn) p[n] = rv[i] + (T)beta * p[n];
        }
        __syncthreads();
        gamma = rr[0];
        ++k;
    }
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int n = threadIdx.x + i * bs;
        if (n < nn) x_out[n] = (pin && pin[n]) ? T(0) : xr[i];
    }
    if (threadIdx.x == 0) st->coarse_k = k;
}

// Pinned, coarse grid of at most 64 * NPT nodes: one wavefront, no block barriers and no
// branches in the iteration.  Each owned node keeps its assembled 9-point row of L (boundary,
// missing-element and pinned-row masks folded into the weights) and its clamped neighbour ids in
// registers; the search direction lives in LDS (idle lanes store to a scratch slot past the end);
// the dot products are DPP wavefront sums.
template <class T, class TR, int NPT>
__global__ __launch_bounds__(64) void k_coarse_cg_wave(Grid2D g, const ElemConst* __restrict__ ec,
                                                       const uint8_t* pin, const TR* r, T* x_out, DevState* st,
                                                       double tol, int maxit, const int* active) {
    if (active && *active == 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* p = reinterpret_cast<T*>(smem_raw);  // nn + 64 entries
    const int nn = g.nn, lane = threadIdx.x;
    T w[NPT][9], xr[NPT], rv[NPT], m[NPT];
    int id[NPT][9], slot[NPT];
    double g2 = 0.0;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int n = lane + 64 * i;
        const bool own = n < nn;
        const int nc = own ? n : 0;
        NodeNbr s;
        node_nbr(g, nc, s);
        const bool pinned = pin && pin[nc];
        double wd[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) wd[q] = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (s.has[k]) wd[kCy(k, b) * 3 + kCx(k, b)] += ec->Le[kRow(k) * 4 + b];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            w[i][q] = (own && !pinned) ? (T)wd[q] : T(0);
            id[i][q] = s.nb[q / 3][q % 3];
        }
        slot[i] = own ? n : nn + lane;
        m[i] = own ? T(1) : T(0);
        xr[i] = T(0);
        const T b = (own && !pinned) ? (T)r[nc] : T(0);
        rv[i] = b;
        p[slot[i]] = b;
        g2 += (double)b * (double)b;
    }
    __syncthreads();
    double gamma = wave_sum_dpp(g2);
    const double atol2 = fmax(tol * tol * gamma, 0.0);
    int k = 0;
    while (gamma > atol2 && k < maxit) {
        T ap[NPT], pv[NPT];
        double pap = 0.0;
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            T a = T(0);
#pragma unroll
            for (int q = 0; q < 9; ++q) a += w[i][q] * p[id[i][q]];
            ap[i] = a;
            pv[i] = p[slot[i]];
            pap += (double)m[i] * (double)pv[i] * (double)a;
        }
        pap = wave_sum_dpp(pap);
        const double alpha = gamma / pap;
        double rr = 0.0;
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            xr[i] = xr[i] + (T)alpha * pv[i];
            rv[i] = rv[i] - (T)alpha * ap[i];
            rr += (double)m[i] * (double)rv[i] * (double)rv[i];
        }
        rr = wave_sum_dpp(rr);
        const double beta = rr / gamma;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NPT; ++i) p[slot[i]] = rv[i] + (T)beta * pv[i];
        __syncthreads();
        gamma = rr;
        ++k;
    }
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int n = lane + 64 * i;
        if (n < nn) x_out[n] = (pin && pin[n]) ? T(0) : xr[i];
    }
    if (lane == 0) st->coarse_k = k;
}

// Pinned, larger coarse grids: same recurrence on global (L2-resident) vectors.
template <class T, class TR>
__global__ __launch_bounds__(1024) void k_coarse_cg_global(Grid2D g, const ElemConst* __restrict__ ec,
                                                           const uint8_t* pin, const TR* r, T* x, T* cr, T* cp,
                                                           T* cAp, DevState* st, double tol, int maxit,
                                                           const int* active) {
    if (active && *active == 0) return;
    const int nn = g.nn;
    double g2[1] = {0.0};
    for (int n = threadIdx.x; n < nn; n += blockDim.x) {
        const T b = (pin && pin[n]) ? T(0) : (T)r[n];
        cr[n] = b; cp[n] = b; x[n] = T(0);
        g2[0] += (double)b * (double)b;
    }
    block_sum<1>(g2);
    double gamma = g2[0];
    const double atol2 = fmax(tol * tol * gamma, 0.0);
    int k = 0;
    while (gamma > atol2 && k < maxit) {
        double pap[1] = {0.0};
        for (int n = threadIdx.x; n < nn; n += blockDim.x) {
            NodeNbr s;
            node_nbr(g, n, s);
            T X[3][3];
            load3x3(s, [&](int m) { return cp[m]; }, X);
            const T a = (pin && pin[n]) ? T(0) : ebe_const(s, ec->Le, X);
            cAp[n] = a;
            pap[0] += (double)X[1][1] * (double)a;
        }
        block_sum<1>(pap);
        const double alpha = gamma / pap[0];
        double rr[1] = {0.0};
        for (int n = threadIdx.x; n < nn; n += blockDim.x) {
            x[n] = x[n] + (T)alpha * cp[n];
            const T rn = cr[n] - (T)alpha * cAp[n];
            cr[n] = rn;
            rr[0] += (double)rn * (double)rn;
        }
        block_sum<1>(rr);
        const double beta = rr[0] / gamma;
        for (int n = threadIdx.x; n < nn; n += blockDim.x) cp[n] = cr[n] + (T)beta * cp[n];
        __syncthreads();
        gamma = rr[0];
        ++k;
    }
    for (int n = threadIdx.x; n < nn; n += blockDim.x)
        if (pin && pin[n]) x[n] = T(0);
    if (threadIdx.x == 0) st->coarse_k = k;
}

template <class T, class TR>
static void coarse_solve(Ctx& c, const TR* r, T* y, T* cr, T* cp, T* cAp, double tol, const int* active) {
    const int l = c.nlev - 1;
    Level& C = c.lv[l];
    const int nn = C.g.nn;
    if (!c.set_zeroP) {
        const T* Ld = std::is_same<T, float>::value ? (const T*)C.Ld32 : (const T*)C.Ld;
        k_coarse_jacobi<T, TR><<<1, 1024, 0, c.stream>>>(nn, Ld, r, y, c.st, l, active);
    } else if (nn <= 64 * 16) {
        const size_t shm = (size_t)(nn + 64) * sizeof(T);
        const int w = (nn + 63) / 64;
        if (w <= 1)
            k_coarse_cg_wave<T, TR, 1><<<1, 64, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
        else if (w <= 2)
            k_coarse_cg_wave<T, TR, 2><<<1, 64, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
        else if (w <= 4)
            k_coarse_cg_wave<T, TR, 4><<<1, 64, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
        else if (w <= 6)
            k_coarse_cg_wave<T, TR, 6><<<1, 64, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
        else if (w <= 8)
            k_coarse_cg_wave<T, TR, 8><<<1, 64, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
        else
            k_coarse_cg_wave<T, TR, 16><<<1, 64, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
    } else if (nn <= 4096) {
        const int npt = nn <= 1024 ? 1 : (nn <= 2048 ? 2 : 4);
        int bs = (nn + npt - 1) / npt;
        bs = (bs + 63) / 64 * 64;
        const size_t shm = (size_t)nn * sizeof(T);
        if (npt == 1)
            k_coarse_cg_lds<T, TR, 1><<<1, bs, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
        else if (npt == 2)
            k_coarse_cg_lds<T, TR, 2><<<1, bs, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
        else
            k_coarse_cg_lds<T, TR, 4><<<1, bs, shm, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, c.st, tol, 1000, active);
    } else {
        k_coarse_cg_global<T, TR><<<1, 1024, 0, c.stream>>>(C.g, C.ec_dev, C.pin, r, y, cr, cp, cAp, c.st, tol,
                                                            1000, active);
    }
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ dense coarse solve
// Mixed form, pinned pressure: y = Ainv b with b = r zeroed at the pins, Ainv the exact inverse
// of the pinned coarse operator (FP32 entries, FP64 accumulation).  Each wave computes
// DENSE_ROWS rows: its lanes stride the columns (independent coalesced loads), then a DPP wave
// sum per row.  ceil(nn / (4 DENSE_ROWS)) blocks of 4 waves.
constexpr int DENSE_ROWS = 2;
template <class T, bool EARLY = false>
__global__ __launch_bounds__(256) void k_coarse_dense(int nn, const float* __restrict__ Ainv,
                                                      const uint8_t* __restrict__ pin, const T* __restrict__ r,
                                                      T* __restrict__ y, const int* active) {
    if constexpr (EARLY) {
        if (*active == 0) return;
    }
    const int go = gate_read(active);
    const int lane = threadIdx.x & 63;
    const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * DENSE_ROWS;
    double acc[DENSE_ROWS];
#pragma unroll
    for (int m = 0; m < DENSE_ROWS; ++m) acc[m] = 0.0;
#pragma unroll 6
    for (int j = lane; j < nn; j += 64) {
        const double bj = (pin && pin[j]) ? 0.0 : (double)r[j];
#pragma unroll
        for (int m = 0; m < DENSE_ROWS; ++m) {
            const int i = row0 + m < nn ? row0 + m : nn - 1;
            acc[m] += (double)Ainv[(size_t)i * nn + j] * bj;
        }
    }
#pragma unroll
    for (int m = 0; m < DENSE_ROWS; ++m) {
        const double v = wave_sum_dpp(acc[m]);
        const int i = row0 + m;
        if (gate_open(go) && lane == 0 && i < nn) y[i] = (pin && pin[i]) ? T(0) : (T)v;
    }
}

void coarse_dense64(Ctx& c, int nn, const uint8_t* pin, const double* r, double* y, const int* active) {
    const int nb = (nn + 4 * DENSE_ROWS - 1) / (4 * DENSE_ROWS);
    k_coarse_dense<double><<<nb, 256, 0, c.stream>>>(nn, c.coarse_inv32, pin, r, y, active);
    FCFD_CHECK_LAUNCH();
}

template <class T>
static LdTab<T> ld_tab(const Level& L) {
    LdTab<T> t;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) t.v[a][b] = (T)L.ldc[a][b];
    return t;
}

// The tiled V-cycle can carry the PCG's x / r update (first down leg) and z-step (last up leg)
// when both fine-level legs run as single-level kernels.
// y = 2/3 r / L_diag: one damped Jacobi from zero with no pinned rows
template <class T>
__global__ __launch_bounds__(256) void k_coarse_jacobi0(int nn, const T* __restrict__ Ld, const T* __restrict__ r,
                                                        T* __restrict__ y, const int* __restrict__ active) {
    if (active && *active == 0) return;
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n < nn) y[n] = (T(OMEGA) * r[n]) / Ld[n];
}

static bool cg_fusable(const Ctx& c) {
    if (!(c.precond == FCFD_PC_MULTIGRID && (c.set_zeroP || c.mg_neu_tiled) && !c.g.periodic && c.nlev >= 2 &&
          c.mg_tiled && c.c_r2))
        return false;
    const int ld = c.nlev - 1;
    const bool down_single = !(1 < ld && c.lv[1].g.nny < 200);
    const bool up_single = c.lv[0].g.nn >= (1 << 17);  // never paired with level 1 (vcycle_tiled's up leg)
    return down_single && up_single;
}

// Whether the PCG's z buffers hold FP32 values: the mixed form on the fused tiled iteration,
// whose z-step (the tiled FP32 V-cycle's last up leg) stores z as the FP32 y it is and whose K1
// widens it on load -- the FP64 values an FP64 store of the widened y would hold.
static bool z_f32(const Ctx& c) { return c.precision == FCFD_PREC_MIXED && cg_fusable(c); }

// The pinned V-cycle on a non-periodic grid with the tiled kernels (fcfd_mgtile.hpp): two small
// levels per kernel on each leg below 200 node rows (down) / 2^17 nodes (up), one level per
// kernel above; on the up leg, three levels in one kernel where a pair would leave the level
// below it (at least 2^17 nodes) to a kernel of its own.  With `tail` the level-0 up leg (when it runs alone) also runs the PCG's z-step;
// the return value says so.
template <class T>
static bool vcycle_tiled(Ctx& c, const double* r0, const int* active, const CgTail* tail) {
    if (tail && tail->cg_p && !cg_fusable(c)) throw Error(FCFD_ERR_STATE, "CG update not fusable here");
    using B = LvBufs<T>;
    const int N = c.nlev, ld = N - 1;
    // pure Neumann (mixed form): no pinned rows anywhere
    auto pn = [&](const Level& X) -> const uint8_t* { return X.pin ? X.pin : c.zero_pin; };
    bool fused = false;
    int down_nb = 0;  // blocks of the first down leg when it carries the CG's x / r update
    for (int l = 0; l < ld;) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        if (l + 1 < ld && Lc.g.nny < 200) {
            Level& L2 = c.lv[l + 2];
            const dim3 grid((L2.g.nnx + M2_C - 1) / M2_C, (L2.g.nny + M2_R - 1) / M2_R);
            with_gate_mode(c, active, [&](auto e) {
                constexpr bool E = decltype(e)::value;
                if (l == 0)
                    k_mgt_down2<T, double, E><<<grid, 256, 0, c.stream>>>(
                        L.g, Lc.g, L2.g, L.ec_dev->Le, Lc.ec_dev->Le, ld_tab<T>(L), ld_tab<T>(Lc), r0, pn(L), pn(Lc),
                        B::r(Lc), B::r(L2), active);
                else
                    k_mgt_down2<T, T, E><<<grid, 256, 0, c.stream>>>(
                        L.g, Lc.g, L2.g, L.ec_dev->Le, Lc.ec_dev->Le, ld_tab<T>(L), ld_tab<T>(Lc), B::r(L), pn(L),
                        pn(Lc), B::r(Lc), B::r(L2), active);
            });
            l += 2;
        } else {
            auto go = [&](auto rl) {
                using TR = std::remove_const_t<std::remove_pointer_t<decltype(rl)>>;
                constexpr int CX = MGT_DOWN_CX, CY = MGT_DOWN_CY;
                const int nbx = (Lc.g.nnx + CX - 1) / CX;
                with_gate_mode(c, active, [&](auto e) {
                    k_mgt_down<T, TR, CX, CY, decltype(e)::value>
                        <<<dim3(nbx, (Lc.g.nny + CY - 1) / CY), 256, 0, c.stream>>>(
                            L.g, Lc.g, L.ec_dev->Le, ld_tab<T>(L), rl, pn(L), B::r(Lc), active);
                });
            };
            if (l == 0 && tail && tail->cg_p) {
                constexpr int CX = MGT_DOWN_CX, CY = MGT_DOWN_CY;
                const dim3 grid((Lc.g.nnx + CX - 1) / CX, (Lc.g.nny + CY - 1) / CY);
                down_nb = (int)(grid.x * grid.y);
                if (down_nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
                with_gate_mode(c, active, [&](auto e) {
                    k_mgt_down_cg<T, CX, CY, decltype(e)::value><<<grid, 256, 0, c.stream>>>(
                        L.g, Lc.g, L.ec_dev->Le, ld_tab<T>(L), r0, tail->cg_Ap, tail->cg_p, tail->cg_x,
                        tail->cg_rout, pn(L), B::r(Lc), c.st, c.partials, c.pstride, 2, active);
                });
                r0 = tail->cg_rout;
            } else if (l == 0) {
                go(r0);
            } else {
                go((const T*)B::r(L));
            }
            l += 1;
        }
        FCFD_CHECK_LAUNCH();
    }
    {
        Level& C = c.lv[N - 1];
        const bool f32 = std::is_same<T, float>::value;
        const double ctol = f32 ? (c.tol > 1e-5 ? c.tol : 1e-5) : c.tol;
        if (!c.set_zeroP) {
            // the reference's pure-Neumann coarse solve: one Jacobi from zero (without the mean
            // projections, which this V-cycle leaves out)
            k_coarse_jacobi0<T><<<(C.g.nn + 255) / 256, 256, 0, c.stream>>>(C.g.nn, B::Ld(C), B::r(C), B::y(C),
                                                                              active);
            FCFD_CHECK_LAUNCH();
        } else if (f32 && c.coarse_inv32) {
            const int nb = (C.g.nn + 4 * DENSE_ROWS - 1) / (4 * DENSE_ROWS);
            with_gate_mode(c, active, [&](auto e) {
                k_coarse_dense<T, decltype(e)::value><<<nb, 256, 0, c.stream>>>(C.g.nn, c.coarse_inv32, pn(C),
                                                                                 B::r(C), B::y(C), active);
            });
            FCFD_CHECK_LAUNCH();
        } else {
            coarse_solve<T, T>(c, B::r(C), B::y(C), B::cr(c), B::cp(c), B::cAp(c), ctol, active);
        }
    }
    for (int l = ld - 1; l >= 0;) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        if (l >= 3 && c.lv[l - 1].g.nn < (1 << 17) && c.lv[l - 2].g.nn >= (1 << 17)) {
            // levels l, l - 1 and l - 2 in one kernel, where a pair (l, l - 1) would leave level
            // l - 2 to a one-level kernel of its own (1024^2: levels 3, 2, 1)
            Level& Lm = c.lv[l - 1];
            Level& Lf = c.lv[l - 2];
            const dim3 grid((Lf.g.nnx + U3_C - 1) / U3_C, (Lf.g.nny + U3_R - 1) / U3_R);
            with_gate_mode(c, active, [&](auto e) {
                k_mgt_up3<T, decltype(e)::value><<<grid, 256, 0, c.stream>>>(
                    Lf.g, Lm.g, L.g, Lc.g, Lf.ec_dev->Leoff, Lm.ec_dev->Leoff, L.ec_dev->Leoff, ld_tab<T>(Lf),
                    ld_tab<T>(Lm), ld_tab<T>(L), B::r(Lf), pn(Lf), B::r(Lm), pn(Lm), B::r(L), pn(L), B::y(Lc),
                    B::y(Lf), active);
            });
            FCFD_CHECK_LAUNCH();
            l -= 3;
            continue;
        }
        if (l >= 1 && c.lv[l - 1].g.nn < (1 << 17)) {
            Level& Lf = c.lv[l - 1];
            const dim3 grid((Lf.g.nnx + U2_C - 1) / U2_C, (Lf.g.nny + U2_R - 1) / U2_R);
            with_gate_mode(c, active, [&](auto e) {
                constexpr bool E = decltype(e)::value;
                if (l - 1 == 0)
                    k_mgt_up2<T, double, E><<<grid, 256, 0, c.stream>>>(
                        Lf.g, L.g, Lc.g, Lf.ec_dev->Leoff, L.ec_dev->Leoff, ld_tab<T>(Lf), ld_tab<T>(L), r0, pn(Lf),
                        B::r(L), pn(L), B::y(Lc), B::y(Lf), active);
                else
                    k_mgt_up2<T, T, E><<<grid, 256, 0, c.stream>>>(
                        Lf.g, L.g, Lc.g, Lf.ec_dev->Leoff, L.ec_dev->Leoff, ld_tab<T>(Lf), ld_tab<T>(L), B::r(Lf),
                        pn(Lf), B::r(L), pn(L), B::y(Lc), B::y(Lf), active);
            });
            FCFD_CHECK_LAUNCH();
            l -= 2;
            continue;
        }
        // one level: tiles of MGT_UP_C x MGT_UP_R nodes
        auto go = [&](auto rl, auto zm, auto zt, double* znew, const double* zold, double* pzero, int flex) {
            using TR = std::remove_const_t<std::remove_pointer_t<decltype(rl)>>;
            using ZT = typename decltype(zt)::type;
            constexpr int FC = MGT_UP_C, FR = MGT_UP_R, ZM = decltype(zm)::value;
            const dim3 grid((L.g.nnx + FC - 1) / FC, (L.g.nny + FR - 1) / FR);
            const int nb = (int)(grid.x * grid.y);
            if (ZM != 0 && nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
            with_gate_mode(c, active, [&](auto e) {
                k_mgt_up<T, TR, FC, FR, ZM, ZT, decltype(e)::value><<<grid, 256, 0, c.stream>>>(
                    L.g, Lc.g, L.ec_dev->Leoff, ld_tab<T>(L), rl, pn(L), B::y(Lc), B::y(L), (ZT*)znew,
                    (const ZT*)zold, pzero, flex, c.partials, c.pstride, active);
            });
            if constexpr (ZM == 1)
                finalize_kernel<1><<<1, FIN_BLOCK, 0, c.stream>>>(nb, active, CgGamma0Fin{c.st}, c.partials,
                                                                  c.pstride);
            if constexpr (ZM == 2) {
                const ChunkEnd ce = tail->ce;
                const int* fa = ce.rec ? nullptr : active;
                if (down_nb > 0)
                    finalize2_kernel<2, 1><<<1, FIN_BLOCK, 0, c.stream>>>(nb, down_nb, fa, CgK3RFin{c.st, flex, ce},
                                                                          c.partials, c.pstride);
                else
                    finalize_kernel<2><<<1, FIN_BLOCK, 0, c.stream>>>(nb, fa, CgK3Fin{c.st, flex, ce}, c.partials,
                                                                      c.pstride);
            }
        };
        using Z0 = std::integral_constant<int, 0>;
        using Z1 = std::integral_constant<int, 1>;
        using Z2 = std::integral_constant<int, 2>;
        using ZF = type_tag<float>;
        using ZD = type_tag<double>;
        const bool z32 = std::is_same<T, float>::value && tail && tail->z32;
        if (l == 0 && tail && tail->first) {
            if (z32) go(r0, Z1{}, ZF{}, tail->znew, nullptr, tail->pzero, 0);
            else go(r0, Z1{}, ZD{}, tail->znew, nullptr, tail->pzero, 0);
            fused = true;
        } else if (l == 0 && tail) {
            if (z32) go(r0, Z2{}, ZF{}, tail->znew, tail->zold, nullptr, tail->flex);
            else go(r0, Z2{}, ZD{}, tail->znew, tail->zold, nullptr, tail->flex);
            fused = true;
        } else if (l == 0) {
            go(r0, Z0{}, type_tag<T>{}, nullptr, nullptr, nullptr, 0);
        } else {
            go((const T*)B::r(L), Z0{}, type_tag<T>{}, nullptr, nullptr, nullptr, 0);
        }
        FCFD_CHECK_LAUNCH();
        l -= 1;
    }
    return fused;
}

// ------------------------------------------------------------------------------ V-cycle driver
// With `tail` (pinned form) the level-0 up leg, when staged, also runs the PCG's z-step and the
// return value says so (nothing is then stored into the level-0 y).
template <class T>
static bool vcycle_t(Ctx& c, const double* r0, const int* active, const CgTail* tail = nullptr) {
    if ((c.set_zeroP || (std::is_same<T, float>::value && c.mg_neu_tiled)) && !c.g.periodic && c.nlev >= 2 &&
        c.mg_tiled)
        return vcycle_tiled<T>(c, r0, active, tail);
    bool fused = false;
    using B = LvBufs<T>;
    const int N = c.nlev;
    const int neu = !c.set_zeroP;
    const bool f32 = std::is_same<T, float>::value;
    const double ctol = f32 ? (c.tol > 1e-5 ? c.tol : 1e-5) : c.tol;
    const int ld = N - 1;  // levels [0, ld) above the coarse level
    // pinned form: no mean projections, so its kernels carry no reduction (and the small levels
    // run in 64-thread blocks)
    auto run = [&](int n, const auto& body, const MeanFin& fin) {
        if (neu) launch_node<1>(c, n, active, body, fin);
        else if (n < (1 << 17)) launch_node_nored<64>(c, n, active, body);
        else launch_node<0>(c, n, active, body, fin);
    };
    for (int l = 0; l < ld; ++l) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        const int n = L.g.nn;
        if (!neu && !L.g.periodic) {
            if (l + 1 < ld && Lc.g.nny < 200) {
                // two small levels in one kernel
                Level& L2 = c.lv[l + 2];
                const dim3 grid((L2.g.nnx + D2_TC - 1) / D2_TC, (L2.g.nny + D2_TR - 1) / D2_TR);
                if (l == 0)
                    k_down2<T, double><<<grid, 256, 0, c.stream>>>(L.g, Lc.g, L2.g, L.ec_dev->Le, Lc.ec_dev->Le, r0,
                                                                   B::Ld(L), L.pin, B::x(L), B::Ld(Lc), Lc.pin,
                                                                   B::r(Lc), B::x(Lc), B::r(L2), active);
                else
                    k_down2<T, T><<<grid, 256, 0, c.stream>>>(L.g, Lc.g, L2.g, L.ec_dev->Le, Lc.ec_dev->Le, B::r(L),
                                                              B::Ld(L), L.pin, B::x(L), B::Ld(Lc), Lc.pin, B::r(Lc),
                                                              B::x(Lc), B::r(L2), active);
                FCFD_CHECK_LAUNCH();
                ++l;
                continue;
            }
            if (l == 0)
                launch_down_fused<T, double>(c, L.g, Lc.g, L.ec_dev->Le, r0, B::Ld(L), L.pin, B::x(L), B::r(Lc),
                                             active);
            else
                launch_down_fused<T, T>(c, L.g, Lc.g, L.ec_dev->Le, B::r(L), B::Ld(L), L.pin, B::x(L), B::r(Lc),
                                        active);
            FCFD_CHECK_LAUNCH();
            continue;
        }
        if (l == 0) {
            if (neu)
                launch_node<1>(c, n, active, DownX<T, double>{r0, B::Ld(L), B::x(L), c.st, l},
                               MeanFin{&c.st->mg[l][MG_X], nullptr, n});
            if (L.g.periodic)
                run(n, DownRes<T, double>{L.g, L.ec_dev->Le, r0, B::Ld(L), L.pin, B::x(L), B::f(L), c.st, l, neu},
                    MeanFin{&c.st->mg[l][MG_F], nullptr, n});
            else if (neu)
                launch_stencil<1>(c, n, active,
                                  DownResS<T, double>{L.g, L.ec_dev->Le, r0, B::Ld(L), L.pin, B::x(L), B::f(L), c.st,
                                                      l, neu},
                                  MeanFin{&c.st->mg[l][MG_F], nullptr, n});
            else
                launch_stencil<0>(c, n, active,
                                  DownResS<T, double>{L.g, L.ec_dev->Le, r0, B::Ld(L), L.pin, B::x(L), B::f(L), c.st,
                                                      l, neu},
                                  MeanFin{&c.st->mg[l][MG_F], nullptr, n});
        } else {
            if (neu)
                launch_node<1>(c, n, active, DownX<T, T>{B::r(L), B::Ld(L), B::x(L), c.st, l},
                               MeanFin{&c.st->mg[l][MG_X], nullptr, n});
            run(n, DownRes<T, T>{L.g, L.ec_dev->Le, B::r(L), B::Ld(L), L.pin, B::x(L), B::f(L), c.st, l, neu},
                MeanFin{&c.st->mg[l][MG_F], nullptr, n});
        }
        run(Lc.g.nn, Restrict<T>{L.g, Lc.g, B::f(L), B::r(Lc), c.st, l, neu},
            MeanFin{&c.st->mg[l][MG_C], &c.st->mg[l + 1][MG_R], Lc.g.nn});
    }
    {
        Level& C = c.lv[N - 1];
        if (f32 && c.coarse_inv32 && N > 1) {
            const int nb = (C.g.nn + 4 * DENSE_ROWS - 1) / (4 * DENSE_ROWS);
            k_coarse_dense<T><<<nb, 256, 0, c.stream>>>(C.g.nn, c.coarse_inv32, C.pin, B::r(C), B::y(C), active);
            FCFD_CHECK_LAUNCH();
        } else if (N == 1) {
            coarse_solve<T, double>(c, r0, B::y(C), B::cr(c), B::cp(c), B::cAp(c), ctol, active);
        } else {
            coarse_solve<T, T>(c, B::r(C), B::y(C), B::cr(c), B::cp(c), B::cAp(c), ctol, active);
        }
    }
    for (int l = ld - 1; l >= 0; --l) {
        Level& L = c.lv[l];
        Level& Lc = c.lv[l + 1];
        const int n = L.g.nn;
        if (neu)
            launch_node<1>(c, n, active, Prolong<T>{L.g, Lc.g, B::y(Lc), B::f(L), c.st, l},
                           MeanFin{&c.st->mg[l][MG_P], nullptr, n});
        if (!neu && !L.g.periodic && n >= (1 << 17)) {  // staging pays on the large levels only
            const JacobiS<T, double> j0{L.g, Lc.g, L.ec_dev->Leoff, B::Ld(L), L.pin, r0, B::x(L), B::y(Lc), B::y(L)};
            if (l == 0 && tail && tail->first) {
                launch_stencil<1>(c, n, active, JacobiSZ<T, 1>{j0, tail->znew, nullptr, 0, tail->pzero},
                                  CgGamma0Fin{c.st});
                fused = true;
            } else if (l == 0 && tail) {
                launch_stencil<2>(c, n, active, JacobiSZ<T, 2>{j0, tail->znew, tail->zold, tail->flex, nullptr},
                                  CgK3Fin{c.st, tail->flex, tail->ce});
                fused = true;
            } else if (l == 0)
                launch_stencil<0>(c, n, active, j0, NoFin{});
            else
                launch_stencil<0>(c, n, active,
                                  JacobiS<T, T>{L.g, Lc.g, L.ec_dev->Leoff, B::Ld(L), L.pin, B::r(L), B::x(L),
                                                B::y(Lc), B::y(L)},
                                  NoFin{});
            continue;
        }
        if (!neu && !L.g.periodic && l >= 1 && c.lv[l - 1].g.nn < (1 << 17)) {
            // this level and the next finer one in one kernel
            Level& Lf = c.lv[l - 1];
            const dim3 grid((Lf.g.nnx + U2_LC - 1) / U2_LC, (Lf.g.nny + U2_LR - 1) / U2_LR);
            if (l - 1 == 0)
                k_up2<T, double><<<grid, 256, 0, c.stream>>>(Lf.g, L.g, Lc.g, Lf.ec_dev->Leoff, L.ec_dev->Leoff,
                                                             B::Ld(Lf), Lf.pin, r0, B::x(Lf), B::Ld(L), L.pin,
                                                             B::r(L), B::x(L), B::y(Lc), B::y(Lf), active);
            else
                k_up2<T, T><<<grid, 256, 0, c.stream>>>(Lf.g, L.g, Lc.g, Lf.ec_dev->Leoff, L.ec_dev->Leoff,
                                                        B::Ld(Lf), Lf.pin, B::r(Lf), B::x(Lf), B::Ld(L), L.pin,
                                                        B::r(L), B::x(L), B::y(Lc), B::y(Lf), active);
            FCFD_CHECK_LAUNCH();
            --l;
            continue;
        }
        if (l == 0)
            run(n, Jacobi<T, double>{L.g, Lc.g, L.ec_dev->Leoff, B::Ld(L), L.pin, r0, B::x(L), B::f(L), B::y(Lc),
                                     B::y(L), c.st, l, neu},
                MeanFin{&c.st->mg[l][MG_Y], nullptr, n});
        else
            run(n, Jacobi<T, T>{L.g, Lc.g, L.ec_dev->Leoff, B::Ld(L), L.pin, B::r(L), B::x(L), B::f(L), B::y(Lc),
                                B::y(L), c.st, l, neu},
                MeanFin{&c.st->mg[l][MG_Y], nullptr, n});
    }
    FCFD_CHECK_LAUNCH();
    return fused;
}

// z = y0 - mean (Neumann) as FP64
template <class T>
struct ZOut {
    const T* y; double* z; const DevState* st; int neumann;
    __device__ void operator()(int k, double (&)[1]) const {
        z[k] = (double)y[k] - (neumann ? st->mg[0][MG_Y] : 0.0);
    }
};

void vcycle(Ctx& c, const double* r, double* z) {
    FCFD_REQUIRE(c.precond == FCFD_PC_MULTIGRID, "context was not created with the multigrid preconditioner");
    const int nn = c.g.nn;
    // mean of r for the first down-leg projection (pure Neumann)
    struct SumBody {
        const double* r;
        __device__ void operator()(int k, double (&acc)[1]) const { acc[0] += r[k]; }
    };
    launch_node<1>(c, nn, nullptr, SumBody{r},
                                                         MeanFin{&c.st->mg[0][MG_R], nullptr, nn});
    if (c.precision == FCFD_PREC_MIXED) {
        vcycle_t<float>(c, r, nullptr);
        if (c.mg_neu_tiled) {
            // the tiled pure-Neumann V-cycle leaves the projections out: take the mean of y here
            struct SumY {
                const float* y;
                __device__ void operator()(int k, double (&acc)[1]) const { acc[0] += (double)y[k]; }
            };
            launch_node<1>(c, nn, nullptr, SumY{c.lv[0].y32}, MeanFin{&c.st->mg[0][MG_Y], nullptr, nn});
        }
        launch_node<0>(c, nn, nullptr,
                                                             ZOut<float>{c.lv[0].y32, z, c.st, !c.set_zeroP},
                                                             NoFin{});
    } else {
        vcycle_t<double>(c, r, nullptr);
        launch_node<0>(c, nn, nullptr,
                                                             ZOut<double>{c.lv[0].y, z, c.st, !c.set_zeroP},
                                                             NoFin{});
    }
    FCFD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------------------ PCG
struct CgInit {
    Grid2D g; const double* Le; const uint8_t* pin; const double* b; const double* bmean; const double* x0in;
    int zero_pins; double* x; double* r;
    __device__ void operator()(int n, double (&acc)[3]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3];
        if (zero_pins && pin)
            load3x3(s, [&](int k) { const double v = x0in[k]; return pin[k] ? 0.0 : v; }, X);
        else
            load3x3(s, [&](int k) { return x0in[k]; }, X);
        const bool pn = pin && pin[n];
        const double ax = pn ? 0.0 : ebe_const(s, Le, X);
        const double bn = b[n] - (bmean ? *bmean : 0.0);
        const double rn = bn - ax;
        x[n] = X[1][1];
        r[n] = rn;
        acc[0] += bn * bn;
        acc[1] += rn * rn;
        acc[2] += rn;
    }
};


// CgInit with the pinned pressure rhs formed per node (PrhsBody's value, so bitwise the b that
// CgInit would read): one pass instead of the rhs kernel, its finalize and the b store and load
struct CgInitD {
    Grid2D g; const double* Le; const uint8_t* pin; PrhsSpec ps; const double* x0in; int zero_pins;
    double* x; double* r;
    __device__ void operator()(int n, double (&acc)[3]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3], Xu[3][3], Xv[3][3];
        if (zero_pins && pin)
            load3x3(s, [&](int k) { const double v = x0in[k]; return pin[k] ? 0.0 : v; }, X);
        else
            load3x3(s, [&](int k) { return x0in[k]; }, X);
        load3x3(s, [&](int k) { return ps.Us[k]; }, Xu);
        load3x3(s, [&](int k) { return ps.Us[k + g.nn]; }, Xv);
        const bool pn = pin && pin[n];
        const double div = ebe_const_T(s, ps.ec->Gx, Xu) + ebe_const_T(s, ps.ec->Gy, Xv);
        double bn = ps.c * (div - (ps.Fp ? ps.Fp[n] : 0.0));
        if (pn) bn = 0.0;
        const double ax = pn ? 0.0 : ebe_const(s, Le, X);
        const double rn = bn - ax;
        x[n] = X[1][1];
        r[n] = rn;
        acc[0] += bn * bn;
        acc[1] += rn * rn;
        acc[2] += rn;
    }
};

// p = z + beta p_old (on the fly), Ap = L p (pinned rows 0), finalize alpha = gamma / p.Ap
struct CgK1 {
    Grid2D g; const double* Le; const uint8_t* pin; const double* z; const double* pold; double* pnew; double* Ap;
    const DevState* st;
    __device__ void operator()(int n, double (&acc)[1]) const {
        const double beta = st->c_beta;
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3];
        load3x3(s, [&](int k) { return z[k] + beta * pold[k]; }, X);
        const double a = (pin && pin[n]) ? 0.0 : ebe_const(s, Le, X);
        pnew[n] = X[1][1];
        Ap[n] = a;
        acc[0] += X[1][1] * a;
    }
};
// LDS-staged form (non-periodic grids): p staged once per block window
struct CgK1S {
    static constexpr int NV = 1;
    static constexpr int NPT = 1;
    Grid2D g; const double* Le; const uint8_t* pin; const double* z; const double* pold; double* pnew; double* Ap;
    const DevState* st;
    __device__ __forceinline__ int nnx() const { return g.nnx; }
    __device__ __forceinline__ void stage(int id, double (&v)[1]) const { v[0] = z[id] + st->c_beta * pold[id]; }
    __device__ void operator()(int n, const StageView<1, NPT>& sv, double (&acc)[1]) const {
        NodeNbr s;
        node_elems(g, n, s);
        double X[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) X[a][b] = sv.at(0, n, a - 1, b - 1);
        const double a = (pin && pin[n]) ? 0.0 : ebe_const(s, Le, X);
        pnew[n] = X[1][1];
        Ap[n] = a;
        acc[0] += X[1][1] * a;
    }
};




// Approximate inverse, two-term Neumann series (standard.py:225-234):
//   D1x = r / L_diag;  D1L = L D1x / L_diag, 0 at gamma_p;  z = 3 D1x - 3 D1L + (L D1L, 0 at gamma_p) / L_diag
struct ApproxInv1 {
    Grid2D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; double* d1l; int keep;
    __device__ void operator()(int n, double (&)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3];
        load3x3(s, [&](int k) { return r[k] / Ld[k]; }, X);
        const double v = ebe_const(s, Le, X) / Ld[n];
        d1l[n] = (pin && pin[n]) ? (keep ? X[1][1] : 0.0) : v;
    }
};
struct ApproxInv2 {
    Grid2D g; const double* Le; const double* r; const double* Ld; const uint8_t* pin; const double* d1l; double* z;
    int keep;
    __device__ void operator()(int n, double (&)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3];
        load3x3(s, [&](int k) { return d1l[k]; }, X);
        double l2 = ebe_const(s, Le, X);
        if (pin && pin[n]) l2 = keep ? X[1][1] : 0.0;
        const double d1x = r[n] / Ld[n];
        z[n] = (3.0 * d1x - 3.0 * X[1][1]) + l2 / Ld[n];
    }
};

static bool vcycle_t_mixed_or_64(Ctx& c, const double* r, const int* active, const CgTail* tail) {
    if (c.precision == FCFD_PREC_MIXED) return vcycle_t<float>(c, r, active, tail);
    return vcycle_t<double>(c, r, active, tail);
}

// applies the preconditioner to c.c_r; returns pointer + type of y0.  With `tail`, a V-cycle
// may also run the CG's z-step (*fused = true; y0 is then not written).
static const void* precondition(Ctx& c, const int* active, bool& is_f32, const CgTail* tail = nullptr,
                                bool* fused = nullptr) {
    if (fused) *fused = false;
    const int nn = c.g.nn;
    is_f32 = false;
    switch (c.precond) {
        case FCFD_PC_NONE:
            return c.c_r;
        case FCFD_PC_JACOBI:
            launch_node<0>(c, nn, active, JacobiPrec{c.c_r, c.lv[0].Ld, c.lv[0].y},
                                                                 NoFin{});
            return c.lv[0].y;
        case FCFD_PC_APPROX_INV: {
            const double* Le = c.lv[0].ec_dev->Le;
            launch_node<0>(c, nn, active, ApproxInv1{c.g, Le, c.c_r, c.lv[0].Ld, c.ppin, c.lv[0].x, c.approx_keep},
                           NoFin{});
            launch_node<0>(c, nn, active,
                           ApproxInv2{c.g, Le, c.c_r, c.lv[0].Ld, c.ppin, c.lv[0].x, c.lv[0].y, c.approx_keep},
                           NoFin{});
            return c.lv[0].y;
        }
        case FCFD_PC_MULTIGRID:
            if (c.precision == FCFD_PREC_MIXED) {
                const bool f = vcycle_t<float>(c, c.c_r, active, tail);
                if (fused) *fused = f;
                is_f32 = true;
                return c.lv[0].y32;
            }
            {
                const bool f = vcycle_t<double>(c, c.c_r, active, tail);
                if (fused) *fused = f;
            }
            return c.lv[0].y;
    }
    throw Error(FCFD_ERR_UNSUPPORTED, "pressure preconditioner not supported");
}

// PCG: returns iterations; x0 in x0in (pins zeroed if zero_pins), solution in c.c_x
// one PCG iteration on the p / z double buffers of parity o (the solution accumulates in c.c_x)
static void pcg_one(Ctx& c, int o, ChunkEnd ce = ChunkEnd{}) {
    const Grid2D& g = c.g;
    const int nn = g.nn;
    const int neu = (c.precond == FCFD_PC_MULTIGRID) && !c.set_zeroP;
    const int flex = c.precision == FCFD_PREC_MIXED && c.precond == FCFD_PC_MULTIGRID;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    const double* Le = c.lv[0].ec_dev->Le;
    const int* act = &c.st->c_active;
    const int nw = o ^ 1;
    if (cg_fusable(c)) {
        if (!pin) pin = c.zero_pin;  // pure Neumann: no pinned rows
        // r alternates between c_r (parity 0) and c_r2; the tiled K1, then the V-cycle with the
        // x / r update on its first down leg and the z-step on its last up leg
        double* rin = o == 0 ? c.c_r : c.c_r2;
        double* rout = o == 0 ? c.c_r2 : c.c_r;
        {
            constexpr int FC = CGT_C, FR = CGT_R;
            const dim3 grid((g.nnx + FC - 1) / FC, (g.nny + FR - 1) / FR);
            const int nb = (int)(grid.x * grid.y);
            if (nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
            with_gate_mode(c, act, [&](auto e) {
                constexpr bool E = decltype(e)::value;
                if (z_f32(c))
                    k_cg_k1t<FC, FR, float, E><<<grid, 256, 0, c.stream>>>(g, Le, pin, (const float*)c.c_z[o],
                                                                           c.c_p[o], c.c_p[nw], c.c_Ap, c.st,
                                                                           c.partials, c.pstride, act);
                else
                    k_cg_k1t<FC, FR, double, E><<<grid, 256, 0, c.stream>>>(g, Le, pin, c.c_z[o], c.c_p[o],
                                                                            c.c_p[nw], c.c_Ap, c.st, c.partials,
                                                                            c.pstride, act);
            });
            finalize_kernel<1><<<1, FIN_BLOCK, 0, c.stream>>>(nb, act, CgK1Fin{c.st}, c.partials, c.pstride);
        }
        CgTail tail{c.c_z[nw], c.c_z[o], nullptr, flex, 0};
        tail.cg_p = c.c_p[nw];
        tail.cg_Ap = c.c_Ap;
        tail.cg_x = c.c_x;
        tail.cg_rout = rout;
        tail.ce = ce;
        tail.z32 = z_f32(c);
        const bool f = vcycle_t_mixed_or_64(c, rin, act, &tail);
        if (!f) throw Error(FCFD_ERR_STATE, "fused PCG iteration without its z-step");
        FCFD_CHECK_LAUNCH();
        return;
    }
    if (g.periodic)
        launch_node<1>(c, nn, act, CgK1{g, Le, pin, c.c_z[o], c.c_p[o], c.c_p[nw], c.c_Ap, c.st}, CgK1Fin{c.st});
    else
        launch_stencil<1>(c, nn, act, CgK1S{g, Le, pin, c.c_z[o], c.c_p[o], c.c_p[nw], c.c_Ap, c.st},
                          CgK1Fin{c.st});
    launch_node<2>(c, nn, act, CgK2{c.c_p[nw], c.c_Ap, c.c_x, c.c_r, c.st}, CgK2Fin{c.st, nn});
    bool f, fused;
    CgTail tail{c.c_z[nw], c.c_z[o], nullptr, flex, 0};
    tail.ce = ce;
    const void* y = precondition(c, act, f, neu ? nullptr : &tail, &fused);
    if (!fused && f)
        launch_node<2>(c, nn, act, CgK3<float>{c.c_r, (const float*)y, c.c_z[nw], c.c_z[o], c.st, neu, flex},
                       CgK3Fin{c.st, flex, ce});
    else if (!fused)
        launch_node<2>(c, nn, act, CgK3<double>{c.c_r, (const double*)y, c.c_z[nw], c.c_z[o], c.st, neu, flex},
                       CgK3Fin{c.st, flex, ce});
    FCFD_CHECK_LAUNCH();
}

// n more PCG iterations, the first of parity `launched` & 1 (graph-replayed); with ce.rec the
// last one is launched directly and records the end of a speculative chunk
static void pcg_enqueue(Ctx& c, int launched, int n, ChunkEnd ce = ChunkEnd{}) {
    const double key[4] = {0.0, 0.0, 0.0, 0.0};
    auto one = [&](int o) { pcg_one(c, o); };
    auto pair = [&]() {
        pcg_one(c, 0);
        pcg_one(c, 1);
    };
    if (ce.rec) {
        // iterations past the count the host is sure of test their gates at entry as well
        const int sure = chunk_sure_of(c, n), a = sure < n - 1 ? sure : n - 1;
        enqueue_iters(c, c.g_cg, key, nullptr, launched & 1, a, one, pair);
        {
            EarlyGates eg(c, true);
            enqueue_iters(c, c.g_cg, key, nullptr, (launched + a) & 1, n - 1 - a, one, pair);
        }
        EarlyGates eg(c, sure < n);
        pcg_one(c, (launched + n - 1) & 1, ce);
    } else {
        enqueue_iters(c, c.g_cg, key, nullptr, launched & 1, n, one, pair);
    }
}

// Chunks of PCG iterations until the device flag says done: the first chunk (launched = 0) is as
// long as the previous solve took, later ones add two iterations.
int pcg_resume(Ctx& c, int launched, int maxit, bool* converged) {
    int n = launched > 0 ? 2 : (c.pred_p > 0 ? c.pred_p : 2);
    int h[2] = {0, 0};
    while (true) {
        pcg_enqueue(c, launched, n);
        launched += n;
        n = 2;
        read_state_ints(c, &c.st->c_k, 2, h);  // c_k, c_active
        if (!h[1] || launched >= maxit) break;
    }
    if (converged) *converged = !h[1] && h[0] <= maxit;
    if (h[0] > 0) c.pred_p = h[0];
    return h[0];
}

// U_hat, the lumped mass and dt / rho of a projection U* = U_hat + (dt/rho) G P_n / M that the
// pinned PCG init forms itself (k_cginit_us), storing U* into Us
struct UstarSpec {
    const double* Uh; const double* M; double cus; double* Us;
};

static int pcg_impl(Ctx& c, const double* b, const double* bmean, const double* x0in, int zero_pins, double tol,
                    int maxit, bool* converged, bool zero_exit, const PrhsSpec* prhs, int* chunk, ChunkEnd ce,
                    const UstarSpec* us) {
    const Grid2D& g = c.g;
    const int nn = g.nn;
    const int neu = (c.precond == FCFD_PC_MULTIGRID) && !c.set_zeroP;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    const double* Le = c.lv[0].ec_dev->Le;
    if (us) {
        if (!prhs || !pin || !zero_pins || g.periodic) throw Error(FCFD_ERR_STATE, "fused U* init: pinned form only");
        constexpr int FC = MOMT_C, FR = MOMT_R;
        const dim3 grid((g.nnx + FC - 1) / FC, (g.nny + FR - 1) / FR);
        const int nb = (int)(grid.x * grid.y);
        if (nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
        k_cginit_us<FC, FR><<<grid, 256, 0, c.stream>>>(g, prhs->ec, Le, us->Uh, x0in, us->M, prhs->Fp, pin, us->cus,
                                                        prhs->c, us->Us, c.c_x, c.c_r, c.partials, c.pstride,
                                                        c.gate);
        const CgInitFin fin{c.st, tol, maxit, nn, zero_exit};
        finalize_kernel<3><<<1, FIN_BLOCK, 0, c.stream>>>(nb, fin_gate(c.gate, fin, 0), fin, c.partials,
                                                          c.pstride);
    } else if (prhs)
        launch_node<3>(c, nn, c.gate, CgInitD{g, Le, pin, *prhs, x0in, zero_pins, c.c_x, c.c_r},
                       CgInitFin{c.st, tol, maxit, nn, zero_exit});
    else
        launch_node<3>(c, nn, c.gate, CgInit{g, Le, pin, b, bmean, x0in, zero_pins, c.c_x, c.c_r},
                       CgInitFin{c.st, tol, maxit, nn, zero_exit});
    const int* act = &c.st->c_active;
    bool f32, fused0;
    CgTail tail0{c.c_z[0], nullptr, c.c_p[0], 0, 1};
    tail0.z32 = z_f32(c);
    const void* y0 = precondition(c, act, f32, (neu && !c.mg_neu_tiled) ? nullptr : &tail0, &fused0);
    if (!fused0 && f32)
        launch_node<1>(c, nn, act, CgGamma0<float>{c.c_r, (const float*)y0, c.c_z[0], c.c_p[0], c.st, neu},
                       CgGamma0Fin{c.st});
    else if (!fused0)
        launch_node<1>(c, nn, act, CgGamma0<double>{c.c_r, (const double*)y0, c.c_z[0], c.c_p[0], c.st, neu},
                       CgGamma0Fin{c.st});
    FCFD_CHECK_LAUNCH();
    if (chunk) {
        const int n = *chunk > 0 ? *chunk : (c.pred_p > 0 ? c.pred_p : 2);
        pcg_enqueue(c, 0, n, ce);
        *chunk = n;
        return 0;
    }
    return pcg_resume(c, 0, maxit, converged);
}

int pcg(Ctx& c, const double* b, const double* bmean, const double* x0in, int zero_pins, double tol, int maxit,
        bool* converged, bool zero_exit, const PrhsSpec* prhs, int* chunk, ChunkEnd ce) {
    return pcg_impl(c, b, bmean, x0in, zero_pins, tol, maxit, converged, zero_exit, prhs, chunk, ce, nullptr);
}

// x_out = x - mean(x) + st->mean_rhs (the mean of x0, recorded before the solve)
struct RestoreMeanBody {
    const double* x; double* out; const DevState* st;
    __device__ void operator()(int k, double (&)[1]) const { out[k] = (x[k] - st->mean_p) + st->mean_rhs; }
};

int pressure_cg(Ctx& c, const double* b, double* x, double tol, int maxit, bool* converged) {
    // The tiled pure-Neumann V-cycle (mg_neu_tiled) leaves out the intermediate mean projections,
    // so its z carries a constant and the CG iterate drifts by one.  The reference's iterate keeps
    // the mean of x0 (every search direction has mean zero): the solve records mean(x0) and
    // restores it.  The timestep's pressure output subtracts the mean anyway.
    const int nn = c.g.nn;
    if (c.mg_neu_tiled)
        launch_node<1>(c, nn, nullptr, SumXBody{x}, MeanFin{&c.st->mean_rhs, nullptr, nn});
    int it = pcg(c, b, nullptr, x, 0, tol, maxit, converged);
    if (c.mg_neu_tiled) {
        launch_node<1>(c, nn, nullptr, SumXBody{c.c_x}, MeanFin{&c.st->mean_p, nullptr, nn});
        launch_node<0>(c, nn, nullptr, RestoreMeanBody{c.c_x, x, c.st}, NoFin{});
        FCFD_CHECK_LAUNCH();
        return it;
    }
    FCFD_HIP(hipMemcpyAsync(x, c.c_x, sizeof(double) * nn, hipMemcpyDeviceToDevice, c.stream));
    return it;
}

// ------------------------------------------------------------------------------ solve_pressure
// rhs = rho/dt (D U* - F_p), pinned rows zeroed; sum for the mean projection
struct PrhsBody {
    Grid2D g; const ElemConst* ec; const double* Us; const double* Fp; const uint8_t* pin; double* rhs; double c;
    __device__ void operator()(int n, double (&acc)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double Xu[3][3], Xv[3][3];
        load3x3(s, [&](int k) { return Us[k]; }, Xu);
        load3x3(s, [&](int k) { return Us[k + g.nn]; }, Xv);
        const double div = ebe_const_T(s, ec->Gx, Xu) + ebe_const_T(s, ec->Gy, Xv);
        double v = c * (div - (Fp ? Fp[n] : 0.0));
        if (pin && pin[n]) v = 0.0;
        rhs[n] = v;
        acc[0] += v;
    }
};


// The pressure solve proper: the solution stays in c.c_x (pinned form: 0 at the pins; Neumann:
// its mean is taken by pressure_output).  Reads U_star and P_n (x0) only.
int pressure_solve_only(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                        int* chunk, ChunkEnd ce) {
    const int nn = c.g.nn;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    const int maxit = c.maxit_p > 0 ? c.maxit_p : 1000;
    bool conv;
    int it;
    if (c.set_zeroP) {
        // pinned: no mean to subtract, so the PCG's init kernel forms the rhs itself
        const PrhsSpec ps{c.ec_dev, U_star, F_p, rho / dt};
        it = pcg(c, nullptr, nullptr, P_n, 1, c.tol, maxit, &conv, false, &ps, chunk, ce);
    } else {
        launch_node<1>(c, nn, c.gate, PrhsBody{c.g, c.ec_dev, U_star, F_p, pin, c.c_rhs, rho / dt},
                       MeanFin{&c.st->mean_rhs, nullptr, nn});
        it = pcg(c, c.c_rhs, &c.st->mean_rhs, P_n, 0, c.tol, maxit, &conv, false, nullptr, chunk, ce);
    }
    return it;
}

int ustar_pressure_solve(Ctx& c, const double* U_hat, const double* P_n, const double* F_p, double rho, double dt,
                         int* chunk, ChunkEnd ce) {
    if (c.set_zeroP && !c.g.periodic) {
        const int maxit = c.maxit_p > 0 ? c.maxit_p : 1000;
        bool conv;
        const PrhsSpec ps{c.ec_dev, c.ustar, F_p, rho / dt};
        const UstarSpec us{U_hat, c.Mlump, dt / rho, c.ustar};
        return pcg_impl(c, nullptr, nullptr, P_n, 1, c.tol, maxit, &conv, false, &ps, chunk, ce, &us);
    }
    projection_ustar(c, U_hat, P_n, rho, dt, c.ustar);
    return pressure_solve_only(c, c.ustar, P_n, F_p, rho, dt, chunk, ce);
}

int pressure_resume(Ctx& c, int launched) {
    bool conv;
    return pcg_resume(c, launched, c.maxit_p > 0 ? c.maxit_p : 1000, &conv);
}

// P = the solution with the pins zeroed, or minus its mean (standard.py:119-123)
void pressure_output(Ctx& c, double* P_out) {
    const int nn = c.g.nn;
    const uint8_t* pin = c.set_zeroP ? c.ppin : nullptr;
    if (!c.set_zeroP) launch_node<1>(c, nn, c.gate, SumXBody{c.c_x}, MeanFin{&c.st->mean_p, nullptr, nn});
    launch_node<0>(c, nn, c.gate, PoutBody{c.c_x, pin, P_out, c.st, !c.set_zeroP}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

// The step's last two passes: the pressure output and the final projection from c.ustar, as one
// tiled kernel for the pinned form on non-periodic grids
void output_and_project(Ctx& c, double* P_out, double rho, double dt, double* U) {
    if (c.set_zeroP && !c.g.periodic && c.ppin && c.dmask) {
        const Grid2D& g = c.g;
        constexpr int FC = MOMT_C, FR = MOMT_R;
        const dim3 grid((g.nnx + FC - 1) / FC, (g.nny + FR - 1) / FR);
        k_final_t<FC, FR><<<grid, 256, 0, c.stream>>>(g, c.ec_dev, c.c_x, c.ppin, c.ustar, c.Mlump, c.ud_full,
                                                      c.dmask, dt / rho, P_out, U, c.gate);
        FCFD_CHECK_LAUNCH();
        return;
    }
    pressure_output(c, P_out);
    projection_final(c, c.ustar, P_out, rho, dt, U);
}

int solve_pressure(Ctx& c, const double* U_star, const double* P_n, const double* F_p, double rho, double dt,
                   double* P_out) {
    const int it = pressure_solve_only(c, U_star, P_n, F_p, rho, dt);
    pressure_output(c, P_out);
    return it;
}

// ------------------------------------------------------------------------------ grid operators
// The reference's multigrid building blocks as standalone operators on the levels of a context
// (multigrid.py:90-228): prolongation, restriction and one damped Jacobi sweep.
struct InterpBody {
    Grid2D gf, gc; const double* yc; double* y;
    __device__ void operator()(int n, double (&)[1]) const { y[n] = interp_at(gf, gc, yc, 0.0, n); }
};

void interpolate_level(Ctx& c, int level, const double* xH, double* xh) {
    FCFD_REQUIRE(level >= 0 && level + 1 < c.nlev, "level out of range (the context needs level + 1 too)");
    const Level& L = c.lv[level];
    launch_node<0>(c, L.g.nn, nullptr, InterpBody{L.g, c.lv[level + 1].g, xH, xh}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

void restrict_level(Ctx& c, int level, const double* xh, double* xH) {
    FCFD_REQUIRE(level >= 0 && level + 1 < c.nlev, "level out of range (the context needs level + 1 too)");
    const Level& L = c.lv[level];
    const Level& Lc = c.lv[level + 1];
    launch_node<0>(c, Lc.g.nn, nullptr, Restrict<double>{L.g, Lc.g, xh, xH, c.st, level, 0}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

// x_out = omega (rhs - Le_off x') / L_diag + (1 - omega) x', x' = x with the pins zeroed
// (set_zeroP); then the pins zeroed, or (pure Neumann) the mean of x_out subtracted.
struct ElemMat16 {
    double v[16];
};
struct JacSweepBody {
    Grid2D g; ElemMat16 Lo; const double* x; const double* rhs; const double* Ld; const uint8_t* pin; double* y;
    __device__ void operator()(int n, double (&acc)[1]) const {
        NodeNbr s;
        node_nbr(g, n, s);
        double X[3][3];
        if (pin)
            load3x3(s, [&](int k) { const double v = x[k]; return pin[k] ? 0.0 : v; }, X);
        else
            load3x3(s, [&](int k) { return x[k]; }, X);
        const double off = ebe_const(s, Lo.v, X);
        double v = (OMEGA * (rhs[n] - off)) / Ld[n] + (1.0 - OMEGA) * X[1][1];
        if (pin && pin[n]) v = 0.0;
        y[n] = v;
        acc[0] += v;
    }
};
struct SubMeanBody {
    double* y; const double* mean;
    __device__ void operator()(int n, double (&)[1]) const { y[n] = y[n] - *mean; }
};
__global__ void k_pin_mask(uint8_t* mask, int nn, const int64_t* __restrict__ idx, int64_t n_idx) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nn) mask[k] = 0;
}
__global__ void k_pin_set(uint8_t* mask, int nn, const int64_t* __restrict__ idx, int64_t n_idx) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_idx) return;
    int64_t i = idx[k];
    if (i < 0) i += nn;  // negative indices count from the end, out-of-range ones are dropped (as .at[].set)
    if (i >= 0 && i < nn) mask[i] = 1;
}

void jacobi_sweep(Ctx& c, int level, const double* Le_off16, const double* x, const double* rhs, const double* Ld,
                  const int64_t* gamma_p, int64_t n_gamma_p, int set_zeroP, double* y) {
    FCFD_REQUIRE(level >= 0 && level < c.nlev, "level out of range");
    FCFD_REQUIRE(x != y, "x_out must not alias x (the sweep reads x at the neighbours of every node)");
    const Grid2D& g = c.lv[level].g;
    const int nn = g.nn;
    ElemMat16 Lo;
    for (int k = 0; k < 16; ++k) Lo.v[k] = Le_off16[k];
    const uint8_t* pin = nullptr;
    if (set_zeroP && gamma_p && n_gamma_p > 0) {
        if (!c.sweep_mask) c.sweep_mask = dalloc<uint8_t>(c, c.lv[0].g.nn);
        k_pin_mask<<<(nn + 255) / 256, 256, 0, c.stream>>>(c.sweep_mask, nn, gamma_p, n_gamma_p);
        k_pin_set<<<(int)((n_gamma_p + 255) / 256), 256, 0, c.stream>>>(c.sweep_mask, nn, gamma_p, n_gamma_p);
        pin = c.sweep_mask;
    }
    launch_node<1>(c, nn, nullptr, JacSweepBody{g, Lo, x, rhs, Ld, pin, y}, MeanFin{&c.st->dot, nullptr, nn});
    if (!set_zeroP) launch_node<0>(c, nn, nullptr, SubMeanBody{y, &c.st->dot}, NoFin{});
    FCFD_CHECK_LAUNCH();
}

}  // namespace fcfd
