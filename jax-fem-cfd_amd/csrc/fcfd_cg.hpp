// Geometry-free pieces of the preconditioned CG (jax_fem_cfd/utils/jax_iterative_solvers.py:87-120):
// the finalize steps (device scalar recurrences), the vector-only iteration bodies and the Jacobi
// preconditioner, shared by the 2D and 3D pressure solves.
#pragma once
#include "fcfd_internal.hpp"

namespace fcfd {

struct MeanFin {
    double* slot; double* slot2; int n;
    __device__ void operator()(const double (&t)[1]) const {
        double m = t[0] / n;
        *slot = m;
        if (slot2) *slot2 = m;
    }
};

struct CgInitFin {
    DevState* st; double tol; int maxit; int nn; int zero_exit;
    __device__ void operator()(const double (&t)[3]) const {
        st->c_atol2 = fmax(tol * tol * t[0], 0.0);
        st->c_zero = zero_exit && t[0] == 0.0;
        st->c_rr = t[1];
        st->mg[0][MG_R] = t[2] / nn;
        st->c_k = 0;
        st->c_maxit = maxit;
        st->c_active = (t[1] > st->c_atol2) && (0 < maxit) && !st->c_zero;
    }
};

// z = y0 - mean ; gamma0 = r . z ; p_old = 0 ; beta = 0
template <class T>
struct CgGamma0 {
    const double* r; const T* y; double* z; double* pold; const DevState* st; int neumann;
    __device__ void operator()(int k, double (&acc)[1]) const {
        const double zk = (double)y[k] - (neumann ? st->mg[0][MG_Y] : 0.0);
        z[k] = zk;
        pold[k] = 0.0;
        acc[0] += r[k] * zk;
    }
};

struct CgGamma0Fin {
    DevState* st;
    __device__ void operator()(const double (&t)[1]) const {
        st->c_gamma = t[0];
        st->c_beta = 0.0;
    }
};

struct CgK1Fin {
    DevState* st;
    __device__ void operator()(const double (&t)[1]) const { st->c_alpha = st->c_gamma / t[0]; }
};

// x += alpha p ; r -= alpha Ap ; finalize r.r and mean r
struct CgK2 {
    const double* p; const double* Ap; double* x; double* r; const DevState* st;
    __device__ void operator()(int k, double (&acc)[2]) const {
        const double a = st->c_alpha;
        x[k] = x[k] + a * p[k];
        const double rk = r[k] - a * Ap[k];
        r[k] = rk;
        acc[0] += rk * rk;
        acc[1] += rk;
    }
};

struct CgK2Fin {
    DevState* st; int nn;
    __device__ void operator()(const double (&t)[2]) const {
        st->c_rr = t[0];
        st->mg[0][MG_R] = t[1] / nn;
    }
};

// z_new = y0 - mean ; gamma' = r.z_new ; (flexible: r.z_old) ; beta ; k++ ; convergence flag
template <class T>
struct CgK3 {
    const double* r; const T* y; double* znew; const double* zold; const DevState* st; int neumann; int flex;
    __device__ void operator()(int k, double (&acc)[2]) const {
        const double zk = (double)y[k] - (neumann ? st->mg[0][MG_Y] : 0.0);
        znew[k] = zk;
        acc[0] += r[k] * zk;
        if (flex) acc[1] += r[k] * zold[k];
    }
};

struct CgK3Fin {
    DevState* st; int flex;
    ChunkEnd ce{};  // set on the last iteration of a speculative chunk
    __device__ void update(const double (&t)[2]) const {
        const double gnew = t[0];
        st->c_beta = flex ? (gnew - t[1]) / st->c_gamma : gnew / st->c_gamma;
        st->c_gamma = gnew;
        const int k = st->c_k + 1;
        st->c_k = k;
        st->c_active = (st->c_rr > st->c_atol2) && (k < st->c_maxit);
    }
    __device__ void operator()(const double (&t)[2]) const {
        if (!ce.rec) {
            update(t);
            return;
        }
        if (!st->go) return;
        if (st->c_active) update(t);
        chunk_end(st, ce, st->c_active, st->c_k);
    }
};

// CgK3Fin with the iteration's r . r arriving in the same finalize (t[2]: the x / r update fused
// into the V-cycle's first down leg)
struct CgK3RFin {
    DevState* st; int flex;
    ChunkEnd ce{};
    __device__ void operator()(const double (&t)[3]) const {
        const double t2[2] = {t[0], t[1]};
        if (!ce.rec) {
            st->c_rr = t[2];
            CgK3Fin{st, flex}.update(t2);
            return;
        }
        if (!st->go) return;
        if (st->c_active) {
            st->c_rr = t[2];
            CgK3Fin{st, flex}.update(t2);
        }
        chunk_end(st, ce, st->c_active, st->c_k);
    }
};

struct JacobiPrec {
    const double* r; const double* Ld; double* y;
    __device__ void operator()(int k, double (&)[1]) const { y[k] = r[k] / Ld[k]; }
};

// P = x - mean(x) (Neumann) or x with pins zeroed
struct PoutBody {
    const double* x; const uint8_t* pin; double* P; const DevState* st; int neumann;
    __device__ void operator()(int k, double (&)[1]) const {
        double v = neumann ? x[k] - st->mean_p : x[k];
        if (pin && pin[k]) v = 0.0;
        P[k] = v;
    }
};

struct SumXBody {
    const double* x;
    __device__ void operator()(int k, double (&acc)[1]) const { acc[0] += x[k]; }
};

// x += alpha p ; r -= alpha Ap ; z_new = r / L_diag (jacobi) or r (none) ; partials r.r, r.z_new:
// the CG update, the preconditioner and the new gamma in one pass (CgK2 + JacobiPrec + CgK3)
struct CgK23 {
    const double* p; const double* Ap; const double* Ld; double* x; double* r; double* znew; const DevState* st;
    __device__ void operator()(int k, double (&acc)[2]) const {
        const double a = st->c_alpha;
        x[k] = x[k] + a * p[k];
        const double rk = r[k] - a * Ap[k];
        r[k] = rk;
        const double zk = Ld ? rk / Ld[k] : rk;
        znew[k] = zk;
        acc[0] += rk * rk;
        acc[1] += rk * zk;
    }
};
// CgK2Fin then CgK3Fin (flex = 0)
struct CgK23Fin {
    DevState* st;
    __device__ void operator()(const double (&t)[2]) const {
        st->c_rr = t[0];
        const double gnew = t[1];
        st->c_beta = gnew / st->c_gamma;
        st->c_gamma = gnew;
        const int k = st->c_k + 1;
        st->c_k = k;
        st->c_active = (st->c_rr > st->c_atol2) && (k < st->c_maxit);
    }
};

}  // namespace fcfd
