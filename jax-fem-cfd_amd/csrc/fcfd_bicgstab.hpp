// Generic device-resident BiCGSTAB (jax_fem_cfd/utils/jax_iterative_solvers.py:123-167) over an
// element-by-element operator with NC components sharing one element matrix per element, and
// Dirichlet rows replaced by the identity (standard.py:20-28, passive_transport.py:17-24).
//
// One iteration = 3 kernels (+ their 1-block finalizes):
//   K1  p = r + beta (p - omega q) (formed on the fly at each stencil node, double-buffered),
//       q = A p,  alpha = rho' / (rhat . q).  At the first iteration p0 = q0 = r0 and omega = 1
//       make the bracket exactly 0, so p = r there: K1 reads r alone (b_k == 0) and the init
//       kernel writes no p0, q0 copies of r0.
//   K2  s = r - alpha q (on the fly), t = A s,  exit_early, omega = (t.s)/(t.t)
//   K3  x, r updates,  breakdown codes -10 / -11, next rho', beta, convergence flag
// In the 2D solver s is never stored: K3 forms it again from r and q (bicg_s, one explicit fma,
// so both kernels hold the same value).  The operator apply then writes one vector (t) per
// component instead of two, and K3 reads r and q in place of s.
// Operators provide  double coef(int a, int b, int e)  (row a, column b of element e's matrix).
//
// Masking.  The forward operators replace the Dirichlet rows:  A y = mask ? y : Ahat y.  Their
// transposes (the adjoint solves of implicit differentiation, jax_iterative_solvers.py:209-211)
// mask the input instead:  A^T y = Ahat^T (mask ? 0 : y) + (mask ? y : 0)   (MT = true, with
// TransCoef swapping the element matrix indices).
#pragma once
#include "fcfd_internal.hpp"
#include "fcfd_ops.hpp"

namespace fcfd {

// (cM Me + cL Le + cC Ce_e): the momentum operator and the momentum rhs operator
template <class S>
struct MomCoef {
    Grid2D g;
    const ElemConst* __restrict__ ec;
    S ce;
    double cM, cL, cC;
    __device__ __forceinline__ double coef(int a, int b, int e) const {
        return cM * ec->Me[a * 4 + b] + cL * ec->Le[a * 4 + b] + cC * ce.ld(a * 4 + b, e);
    }
    // row a of element e (one vector load), entries as coef() forms them
    __device__ __forceinline__ void row(int a, int e, double (&cf)[4]) const {
        double v[4];
        ce.ld_row(a, e, v);
#pragma unroll
        for (int b = 0; b < 4; ++b) cf[b] = cM * ec->Me[a * 4 + b] + cL * ec->Le[a * 4 + b] + cC * v[b];
    }
};

// per-element stored matrices (row-blocked): the scalar-transport operators
template <class S>
struct StoredCoef {
    Grid2D g;
    S m;
    __device__ __forceinline__ double coef(int a, int b, int e) const { return m.ld(a * 4 + b, e); }
    __device__ __forceinline__ void row(int a, int e, double (&cf)[4]) const { m.ld_row(a, e, cf); }
};

template <int NC, class Op>
__device__ __forceinline__ void op_apply(const Op& op, const NodeNbr& s, const double (&X)[NC][3][3],
                                         double (&y)[NC]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) y[c] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = kRow(k);
        const int e = s.e[k];
        double t[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) t[c] = 0.0;
        double cf[4];
        op.row(a, e, cf);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
#pragma unroll
            for (int c = 0; c < NC; ++c) t[c] += cf[b] * X[c][kCy(k, b)][kCx(k, b)];
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] += s.has[k] ? t[c] : 0.0;
    }
}

// The init kernel's two operators (rhs and lhs) over the same stencil position.  Generic form:
// two passes.  Two momentum forms of one Ce stream (cM Me + cL Le + cC Ce with different
// coefficients): each Ce row is loaded once and feeds both; every value is formed with the
// operands and operation order of op_apply, so the results are bitwise those of two passes.
template <int NC, class LOp, class ROp>
__device__ __forceinline__ void op_apply_pair(const LOp& lo, const ROp& ro, const NodeNbr& s,
                                              const double (&XL)[NC][3][3], const double (&XR)[NC][3][3],
                                              double (&yl)[NC], double (&yr)[NC]) {
    op_apply<NC>(ro, s, XR, yr);
    op_apply<NC>(lo, s, XL, yl);
}
template <int NC, class S>
__device__ __forceinline__ void op_apply_pair(const MomCoef<S>& lo, const MomCoef<S>& ro, const NodeNbr& s,
                                              const double (&XL)[NC][3][3], const double (&XR)[NC][3][3],
                                              double (&yl)[NC], double (&yr)[NC]) {
    if (lo.ce.c != ro.ce.c) {
        op_apply<NC>(ro, s, XR, yr);
        op_apply<NC>(lo, s, XL, yl);
        return;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) { yl[c] = 0.0; yr[c] = 0.0; }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = kRow(k);
        const int e = s.e[k];
        double v[4];
        lo.ce.ld_row(a, e, v);
        double cl[4], cr[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            cl[b] = lo.cM * lo.ec->Me[a * 4 + b] + lo.cL * lo.ec->Le[a * 4 + b] + lo.cC * v[b];
            cr[b] = ro.cM * ro.ec->Me[a * 4 + b] + ro.cL * ro.ec->Le[a * 4 + b] + ro.cC * v[b];
        }
        double tl[NC], tr[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) { tl[c] = 0.0; tr[c] = 0.0; }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                tl[c] += cl[b] * XL[c][kCy(k, b)][kCx(k, b)];
                tr[c] += cr[b] * XR[c][kCy(k, b)][kCx(k, b)];
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            yl[c] += s.has[k] ? tl[c] : 0.0;
            yr[c] += s.has[k] ? tr[c] : 0.0;
        }
    }
}

// Element matrices transposed (adjoint operators)
template <class Op>
struct TransCoef {
    Op op;
    Grid2D g;
    __device__ __forceinline__ double coef(int a, int b, int e) const { return op.coef(b, a, e); }
    // row a of the transposed matrix = column a of the stored one (4 scalar loads)
    __device__ __forceinline__ void row(int a, int e, double (&cf)[4]) const {
#pragma unroll
        for (int b = 0; b < 4; ++b) cf[b] = op.coef(b, a, e);
    }
};
template <class Op>
__host__ __device__ __forceinline__ TransCoef<Op> transpose_op(const Op& op) { return TransCoef<Op>{op, op.g}; }

// y = masked operator applied to the stencil X (see the header comment for MT)
template <int NC, bool MT, class Op>
__device__ __forceinline__ void masked_apply(const Op& op, const NodeNbr& s, int n, int nn, const uint8_t* mask,
                                             const double (&X)[NC][3][3], double (&y)[NC]) {
    if constexpr (!MT) {
        op_apply<NC>(op, s, X, y);
        if (mask) {
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (mask[n + c * nn]) y[c] = X[c][1][1];
        }
    } else {
        if (mask) {
            double Xm[NC][3][3];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const uint8_t* mc = mask + c * nn;
                uint8_t M[3][3];
                load3x3(s, [&](int k) { return mc[k]; }, M);
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) Xm[c][a][b] = M[a][b] ? 0.0 : X[c][a][b];
            }
            op_apply<NC>(op, s, Xm, y);
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (mask[n + c * nn]) y[c] += X[c][1][1];
        } else {
            op_apply<NC>(op, s, X, y);
        }
    }
}

// No extra rhs term (scalar transport)
struct NoExtra {
    __device__ __forceinline__ void operator()(const NodeNbr&, int, double (&)[1]) const {}
    __device__ __forceinline__ void operator()(const NodeNbr&, int, double (&)[2]) const {}
};

// ------------------------------------------------------------------------------ init
// rhs = R x_n + extra, Dirichlet rows = ud; x0 = x_n with Dirichlet values; r0 = rhs - A x0;
// rhat = p = q = r0.  Reductions: b.b, r0.r0
template <int NC, class LOp, class ROp, class Extra>
struct BInit {
    LOp lhs;
    ROp rhs;
    Extra extra;
    const double *xn, *ud;
    const uint8_t* mask;
    double *x, *r, *rhat;
    __device__ void operator()(int n, double (&acc)[2]) const {
        const int nn = lhs.g.nn;
        NodeNbr s;
        node_nbr(lhs.g, n, s);
        double X[NC][3][3], X0[NC][3][3];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double* xc = xn + c * nn;
            load3x3(s, [&](int k) { return xc[k]; }, X[c]);
            if (mask) {
                const double* uc = ud + c * nn;
                const uint8_t* mc = mask + c * nn;
                load3x3(s, [&](int k) { const double a = uc[k], b = xc[k]; return mc[k] ? a : b; }, X0[c]);
            } else {
                load3x3(s, [&](int k) { return xc[k]; }, X0[c]);
            }
        }
        double b[NC], ax[NC];
        op_apply_pair<NC>(lhs, rhs, s, X0, X, ax, b);
        extra(s, n, b);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const bool m = mask && mask[i];
            const double bc = m ? ud[i] : b[c];
            const double a = m ? X0[c][1][1] : ax[c];
            const double rc = bc - a;
            x[i] = X0[c][1][1];
            r[i] = rc; rhat[i] = rc;
            acc[0] += bc * bc;
            acc[1] += rc * rc;
        }
    }
};
struct BInitFin {
    DevState* st;
    double tol;
    int maxit;
    __device__ void operator()(const double (&t)[2]) const {
        const double a2 = tol * tol * t[0];
        st->b_atol2 = a2 > 0.0 ? a2 : 0.0;
        st->b_rho = 1.0; st->b_alpha = 1.0; st->b_omega = 1.0;
        st->b_rho_new = t[1];
        st->b_beta = t[1] / 1.0 * 1.0 / 1.0;
        st->b_k = 0; st->b_ee = 0; st->b_maxit = maxit;
        st->b_active = (t[1] > st->b_atol2) && (0 < maxit);
    }
};

// ------------------------------------------------------------------------------ K1
template <int NC, class LOp, bool MT = false>
struct BK1 {
    LOp op;
    const uint8_t* mask;
    const double *r, *pold, *qold, *rhat;
    double *pnew, *qnew;
    const DevState* st;
    __device__ void operator()(int n, double (&acc)[1]) const {
        const int nn = op.g.nn;
        const double beta = st->b_beta, omega = st->b_omega;
        NodeNbr s;
        node_nbr(op.g, n, s);
        double X[NC][3][3];
        if (st->b_k == 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const int o = c * nn;
                load3x3(s, [&](int k) { return r[k + o]; }, X[c]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const int o = c * nn;
                load3x3(s, [&](int k) { return r[k + o] + beta * (pold[k + o] - omega * qold[k + o]); }, X[c]);
            }
        }
        double y[NC];
        masked_apply<NC, MT>(op, s, n, nn, mask, X, y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const double yc = y[c];
            pnew[i] = X[c][1][1];
            qnew[i] = yc;
            acc[0] += rhat[i] * yc;
        }
    }
};
struct BK1Fin {
    DevState* st;
    __device__ void operator()(const double (&t)[1]) const { st->b_alpha = st->b_rho_new / t[0]; }
};

// ------------------------------------------------------------------------------ K2
// s = r - alpha q of K2 and K3
__device__ __forceinline__ double bicg_s(double r, double alpha, double q) { return __fma_rn(-alpha, q, r); }

// TAG only distinguishes otherwise identical instantiations (the standalone timing of the
// operator apply runs under its own kernel symbol).
template <int NC, class LOp, int TAG = 0, bool MT = false>
struct BK2 {
    LOp op;
    const uint8_t* mask;
    const double *r, *q;
    double* tv;
    const DevState* st;
    __device__ void operator()(int n, double (&acc)[3]) const {
        const int nn = op.g.nn;
        const double alpha = st->b_alpha;
        NodeNbr s;
        node_nbr(op.g, n, s);
        double X[NC][3][3];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int o = c * nn;
            load3x3(s, [&](int k) { return bicg_s(r[k + o], alpha, q[k + o]); }, X[c]);
        }
        double y[NC];
        masked_apply<NC, MT>(op, s, n, nn, mask, X, y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const double yc = y[c];
            const double sc = X[c][1][1];
            tv[i] = yc;
            acc[0] += sc * sc;
            acc[1] += yc * sc;
            acc[2] += yc * yc;
        }
    }
};
struct BK2Fin {
    DevState* st;
    __device__ void operator()(const double (&t)[3]) const {
        st->b_ee = t[0] < st->b_atol2;
        st->b_omega = t[1] / t[2];
    }
};

// ------------------------------------------------------------------------------ K1 / K2, LDS-staged
// Non-periodic grids: the same recurrences with the on-the-fly vector (p, s) staged per block.
// MT: the staged values are the masked ones (the transposed operator's input); the raw centre
// value comes from global memory.
template <int NC, class LOp, bool MT = false>
struct BK1S {
    static constexpr int NV = NC;
    static constexpr int NPT = 1;
    static constexpr bool LAZY_STAGE = true;
    LOp op;
    const uint8_t* mask;
    const double *r, *pold, *qold, *rhat;
    double *pnew, *qnew;
    const DevState* st;
    __device__ __forceinline__ int nnx() const { return op.g.nnx; }
    __device__ __forceinline__ double pval(int i) const {
        return st->b_k == 0 ? r[i] : r[i] + st->b_beta * (pold[i] - st->b_omega * qold[i]);
    }
    __device__ __forceinline__ void stage(int id, double (&v)[NV]) const {
        const int nn = op.g.nn;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = id + c * nn;
            const double pv = pval(i);
            v[c] = (MT && mask && mask[i]) ? 0.0 : pv;
        }
    }
    __device__ void operator()(int n, const StageView<NV, NPT>& sv, double (&acc)[1]) const {
        const int nn = op.g.nn;
        NodeNbr s;
        node_elems(op.g, n, s);
        double X[NC][3][3];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) X[c][a][b] = sv.at(c, n, a - 1, b - 1);
        double y[NC];
        op_apply<NC>(op, s, X, y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const double pc = MT ? pval(i) : X[c][1][1];
            const bool m = mask && mask[i];
            const double yc = MT ? y[c] + (m ? pc : 0.0) : (m ? pc : y[c]);
            pnew[i] = pc;
            qnew[i] = yc;
            acc[0] += rhat[i] * yc;
        }
    }
};

template <int NC, class LOp, int TAG = 0, bool MT = false>
struct BK2S {
    static constexpr int NV = NC;
    static constexpr int NPT = 1;
    static constexpr bool LAZY_STAGE = true;
    LOp op;
    const uint8_t* mask;
    const double *r, *q;
    double* tv;
    const DevState* st;
    __device__ __forceinline__ int nnx() const { return op.g.nnx; }
    __device__ __forceinline__ double sval(int i) const { return bicg_s(r[i], st->b_alpha, q[i]); }
    __device__ __forceinline__ void stage(int id, double (&v)[NV]) const {
        const int nn = op.g.nn;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = id + c * nn;
            const double s = sval(i);
            v[c] = (MT && mask && mask[i]) ? 0.0 : s;
        }
    }
    __device__ void operator()(int n, const StageView<NV, NPT>& sw, double (&acc)[3]) const {
        const int nn = op.g.nn;
        NodeNbr s;
        node_elems(op.g, n, s);
        double X[NC][3][3];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) X[c][a][b] = sw.at(c, n, a - 1, b - 1);
        double y[NC];
        op_apply<NC>(op, s, X, y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const double sc = MT ? sval(i) : X[c][1][1];
            const bool m = mask && mask[i];
            const double yc = MT ? y[c] + (m ? sc : 0.0) : (m ? sc : y[c]);
            tv[i] = yc;
            acc[0] += sc * sc;
            acc[1] += yc * sc;
            acc[2] += yc * yc;
        }
    }
};

// ------------------------------------------------------------------------------ K3
template <int NC>
struct BK3 {
    int nn;
    const double *p, *sv, *tv, *rhat;
    double *x, *r;
    const DevState* st;
    __device__ void operator()(int k, double (&acc)[2]) const {
        const double alpha = st->b_alpha, omega = st->b_omega;
        const bool ee = st->b_ee;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = k + c * nn;
            const double pi = p[i], si = sv[i], ti = tv[i];
            double xi = x[i];
            double ri;
            if (ee) { xi = xi + alpha * pi; ri = si; }
            else { xi = xi + (alpha * pi + omega * si); ri = si - omega * ti; }
            x[i] = xi; r[i] = ri;
            acc[0] += ri * ri;
            acc[1] += rhat[i] * ri;
        }
    }
};
struct BK3Fin {
    DevState* st;
    ChunkEnd ce{};  // set on the last iteration of a speculative chunk
    __device__ void update(const double (&t)[2]) const {
        const double alpha = st->b_alpha, omega = st->b_omega, rho_cur = st->b_rho_new;
        int k = (omega == 0.0 || alpha == 0.0) ? -11 : st->b_k + 1;
        if (rho_cur == 0.0) k = -10;
        st->b_k = k;
        st->b_rho = rho_cur;
        st->b_rho_new = t[1];
        st->b_beta = t[1] / rho_cur * alpha / omega;
        st->b_active = (t[0] > st->b_atol2) && (k < st->b_maxit) && (k >= 0);
    }
    __device__ void operator()(const double (&t)[2]) const {
        if (!ce.rec) {
            update(t);
            return;
        }
        if (!st->go) return;
        if (st->b_active) update(t);
        chunk_end(st, ce, st->b_active, st->b_k);
    }
};

// K3 of the 2D solver: s formed from r and q (the K2 that precedes it stores no s)
template <int NC>
struct BK3R {
    int nn;
    const double *p, *q, *tv, *rhat;
    double *x, *r;
    const DevState* st;
    __device__ void operator()(int k, double (&acc)[2]) const {
        const double alpha = st->b_alpha, omega = st->b_omega;
        const bool ee = st->b_ee;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = k + c * nn;
            const double pi = p[i], si = bicg_s(r[i], alpha, q[i]), ti = tv[i];
            double xi = x[i];
            double ri;
            if (ee) { xi = xi + alpha * pi; ri = si; }
            else { xi = xi + (alpha * pi + omega * si); ri = si - omega * ti; }
            x[i] = xi; r[i] = ri;
            acc[0] += ri * ri;
            acc[1] += rhat[i] * ri;
        }
    }
};

// ------------------------------------------------------------------------------ init from a vector
// Adjoint solves: rhs = b (a given vector), x0 = x_n with the Dirichlet values (the forward
// solve's x0, which the reference's transposed solve reuses); r0 = b - A x0.  A rhs that is
// exactly zero has the solution 0: the solve is marked done with x = 0 (b_zero) instead of
// iterating towards a zero residual norm.
template <int NC, bool MT, class LOp>
struct BInitVec {
    LOp lhs;
    const double *bvec, *xn, *ud;
    const uint8_t* mask;
    double *x, *r, *rhat;
    __device__ void operator()(int n, double (&acc)[2]) const {
        const int nn = lhs.g.nn;
        NodeNbr s;
        node_nbr(lhs.g, n, s);
        double X0[NC][3][3];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double* xc = xn + c * nn;
            if (mask) {
                const double* uc = ud + c * nn;
                const uint8_t* mc = mask + c * nn;
                load3x3(s, [&](int k) { const double a = uc[k], b = xc[k]; return mc[k] ? a : b; }, X0[c]);
            } else {
                load3x3(s, [&](int k) { return xc[k]; }, X0[c]);
            }
        }
        double ax[NC];
        masked_apply<NC, MT>(lhs, s, n, nn, mask, X0, ax);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int i = n + c * nn;
            const double bc = bvec[i];
            const double rc = bc - ax[c];
            x[i] = X0[c][1][1];
            r[i] = rc; rhat[i] = rc;
            acc[0] += bc * bc;
            acc[1] += rc * rc;
        }
    }
};
struct BInitVecFin {
    DevState* st;
    double tol;
    int maxit;
    __device__ void operator()(const double (&t)[2]) const {
        BInitFin{st, tol, maxit}(t);
        st->b_zero = t[0] == 0.0;
        if (t[0] == 0.0) st->b_active = 0;
    }
};
struct ZeroIfFlag {
    const int* flag;
    double* x;
    __device__ void operator()(int k, double (&)[1]) const {
        if (*flag) x[k] = 0.0;
    }
};

}  // namespace fcfd

#include "fcfd_momtile.hpp"

namespace fcfd {

// ------------------------------------------------------------------------------ driver
// BiCGSTAB iterations after an init kernel has set x, r, rhat, p, q and the state: chunks of
// iteration pairs (graph-replayed) until the device flag says done.  Returns the reference's
// iteration count (or -10 / -11).
// chunk (non-null): only *chunk iterations are enqueued (*chunk <= 0: the predicted count) and no
// state is read; *chunk = iterations enqueued, and a later call with start = that count (after
// the device flag said the solve was still active) completes the solve.
template <int NC, bool MT, class LOp>
int bicgstab_iterate(Ctx& c, IterGraph& graph, const double (&key)[4], const void* kptr, const LOp& lhs,
                     const uint8_t* mask, int maxit, double* x, int& pred, int* chunk = nullptr, int start = 0,
                     ChunkEnd cend = ChunkEnd{}) {
    const int nn = c.g.nn;
    const int* act = &c.st->b_active;
    // one iteration on the double buffers of parity o (ce: the last of a speculative chunk)
    auto one = [&](int o, ChunkEnd ce = ChunkEnd{}) {
        const int nw = o ^ 1;
        if constexpr (NC == 2 && !MT && IsMomCoef<LOp>::value) {
            if (!c.g.periodic && mask) {
                // the tiled momentum kernels (fcfd_momtile.hpp)
                const dim3 grid((c.g.nnx + MOMT_C - 1) / MOMT_C, (c.g.nny + MOMT_R - 1) / MOMT_R);
                const int nb = (int)(grid.x * grid.y);
                if (nb > c.pstride) throw Error(FCFD_ERR_STATE, "reduction partials buffer too small");
                using S = decltype(lhs.ce);
                with_gate_mode(c, act, [&](auto e) {
                    constexpr bool E = decltype(e)::value;
                    k_mom_k1t<S, MOMT_C, MOMT_R, E><<<grid, BLOCK, 0, c.stream>>>(
                        lhs, mask, c.m_r, c.m_p[o], c.m_q[o], c.m_rhat, c.m_p[nw], c.m_q[nw], c.st, c.partials,
                        c.pstride, act);
                    finalize_kernel<1><<<1, FIN_BLOCK, 0, c.stream>>>(nb, act, BK1Fin{c.st}, c.partials, c.pstride);
                    k_mom_k2t<S, MOMT_C, MOMT_R, E ? 2 : 0><<<grid, BLOCK, 0, c.stream>>>(
                        lhs, mask, c.m_r, c.m_q[nw], c.m_t, c.st, c.partials, c.pstride, act);
                });
                finalize_kernel<3><<<1, FIN_BLOCK, 0, c.stream>>>(nb, act, BK2Fin{c.st}, c.partials, c.pstride);
                launch_node<2>(c, nn, act, BK3R<NC>{nn, c.m_p[nw], c.m_q[nw], c.m_t, c.m_rhat, x, c.m_r, c.st},
                               BK3Fin{c.st, ce});
                FCFD_CHECK_LAUNCH();
                return;
            }
        }
        if (c.g.periodic) {
            launch_node<1>(c, nn, act,
                           BK1<NC, LOp, MT>{lhs, mask, c.m_r, c.m_p[o], c.m_q[o], c.m_rhat, c.m_p[nw], c.m_q[nw],
                                            c.st},
                           BK1Fin{c.st});
            launch_node<3>(c, nn, act, BK2<NC, LOp, 0, MT>{lhs, mask, c.m_r, c.m_q[nw], c.m_t, c.st}, BK2Fin{c.st});
        } else {
            launch_stencil<1>(c, nn, act,
                              BK1S<NC, LOp, MT>{lhs, mask, c.m_r, c.m_p[o], c.m_q[o], c.m_rhat, c.m_p[nw],
                                                c.m_q[nw], c.st},
                              BK1Fin{c.st});
            launch_stencil<3>(c, nn, act, BK2S<NC, LOp, 0, MT>{lhs, mask, c.m_r, c.m_q[nw], c.m_t, c.st},
                              BK2Fin{c.st});
        }
        launch_node<2>(c, nn, act, BK3R<NC>{nn, c.m_p[nw], c.m_q[nw], c.m_t, c.m_rhat, x, c.m_r, c.st},
                       BK3Fin{c.st, ce});
        FCFD_CHECK_LAUNCH();
    };
    auto pair = [&]() {
        one(0);
        one(1);
    };
    // the first chunk is as long as the previous solve took; later chunks add two iterations
    int launched = start;
    int n = start > 0 ? 2 : (pred > 0 ? pred : 2);
    int h[2] = {0, 0};
    if (chunk) {
        if (*chunk > 0) n = *chunk;
        if (cend.rec) {
            // the last iteration carries the chunk-end record (launched directly); iterations
            // past the count the host is sure of test their gates at entry as well
            const int sure = chunk_sure_of(c, n), a = sure < n - 1 ? sure : n - 1;
            enqueue_iters(c, graph, key, kptr, launched & 1, a, one, pair);
            {
                EarlyGates eg(c, true);
                enqueue_iters(c, graph, key, kptr, (launched + a) & 1, n - 1 - a, one, pair);
            }
            EarlyGates eg(c, sure < n);
            one((launched + n - 1) & 1, cend);
        } else {
            enqueue_iters(c, graph, key, kptr, launched & 1, n, one, pair);
        }
        *chunk = n;
        return 0;
    }
    while (true) {
        enqueue_iters(c, graph, key, kptr, launched & 1, n, one, pair);
        launched += n;
        read_state_ints(c, &c.st->b_k, 2, h);  // b_k, b_active
        if (!h[1] || launched >= maxit) break;
        n = 2;
    }
    if (h[0] > 0) pred = h[0];
    return h[0];
}

// Solves A x = R x_n + extra (Dirichlet rows: x = ud) from x0 = x_n with Dirichlet values, into
// `x` (NC * nn, a context buffer: the captured iteration graph keeps its address).  Returns the
// reference's iteration count (or -10 / -11).
template <int NC, class LOp, class ROp, class Extra>
int bicgstab_solve(Ctx& c, IterGraph& graph, const double (&key)[4], const void* kptr, const LOp& lhs,
                   const ROp& rhs, const Extra& extra, const double* xn, const double* ud, const uint8_t* mask,
                   int maxit, double* x, int& pred, int* chunk = nullptr, ChunkEnd cend = ChunkEnd{}) {
    launch_node<2>(c, c.g.nn, c.gate,
                   BInit<NC, LOp, ROp, Extra>{lhs, rhs, extra, xn, ud, mask, x, c.m_r, c.m_rhat},
                   BInitFin{c.st, c.tol, maxit});
    FCFD_CHECK_LAUNCH();
    return bicgstab_iterate<NC, false>(c, graph, key, kptr, lhs, mask, maxit, x, pred, chunk, 0, cend);
}

// Transposed solve A^T x = b (adjoint of a masked forward solve), x0 = x_n with Dirichlet values.
template <int NC, class LOp>
int bicgstab_solve_T(Ctx& c, IterGraph& graph, const double (&key)[4], const void* kptr, const LOp& lhsT,
                     const double* b, const double* xn, const double* ud, const uint8_t* mask, int maxit, double* x,
                     int& pred) {
    const int nn = c.g.nn;
    launch_node<2>(c, nn, nullptr,
                   BInitVec<NC, true, LOp>{lhsT, b, xn, ud, mask, x, c.m_r, c.m_rhat},
                   BInitVecFin{c.st, c.tol, maxit});
    launch_node<0>(c, NC * nn, nullptr, ZeroIfFlag{&c.st->b_zero, x}, NoFin{});
    FCFD_CHECK_LAUNCH();
    return bicgstab_iterate<NC, true>(c, graph, key, kptr, lhsT, mask, maxit, x, pred);
}

}  // namespace fcfd
