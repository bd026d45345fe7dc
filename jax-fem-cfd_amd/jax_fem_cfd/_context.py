"""Owning wrapper of one native context (fcfd_ctx) plus the grid registry the
element-operator functions use to find the structured grid behind an
``nconn`` array."""
import ctypes
import math

import numpy as np

from . import _native as nat


class Context:
    """One structured grid + solver configuration on the GPU.

    Mirrors what setup_solver (jax_fem_cfd/solvers/standard.py:191-247) closes
    over: element constants, lumped mass, Dirichlet / pin index sets, the
    multigrid hierarchy and every solver workspace (allocated once here, never
    in the hot calls)."""

    def __init__(self, nex, ney, lx, ly, periodic, *, tol=1e-8, set_zeroP=False, precond=None, mg_levels=1,
                 precision="fp64", ce_format="f64", gamma_d=None, gamma_p=None, mg_gamma_p=None, bconn=None,
                 nvec=None, maxiter_momentum=0, maxiter_pressure=0, gamma_d_phi=None, nez=None, lz=None, order=1,
                 approx_inv_variant=0):
        lib = nat.load_library()
        t = nat.require_gpu()
        self.nex, self.ney, self.lx, self.ly, self.periodic = int(nex), int(ney), float(lx), float(ly), bool(periodic)
        self.ndim = 2 if nez is None else 3
        self.order = int(order)
        k = self.order  # quad9: node grid of 2 nex (+ 1) per direction
        self.nnx = k * self.nex if periodic else k * self.nex + 1
        self.nny = k * self.ney if periodic else k * self.ney + 1
        self.nn = self.nnx * self.nny
        self.ne = self.nex * self.ney
        if self.ndim == 3:
            self.nez, self.lz = int(nez), float(lz)
            self.nnz = k * self.nez + 1  # hex27: node grid of 2 nez + 1
            self.nn *= self.nnz
            self.ne *= self.nez
        self.nen = (3 if self.order == 2 else 2) ** self.ndim  # nodes per element
        self.tol, self.set_zeroP, self.precision, self.ce_format = tol, bool(set_zeroP), precision, ce_format
        self.precond = precond
        md = nat.MeshDesc(self.ndim, self.nex, self.ney, self.lx, self.ly, int(self.periodic),
                          getattr(self, "nez", 0), getattr(self, "lz", 0.0), self.order)
        keep = []

        def arr(a):
            p, k = nat.i64_array(a)
            keep.append(k)
            return p

        sd = nat.SolverDesc()
        sd.tol = float(tol)
        sd.set_zeroP = int(bool(set_zeroP))
        sd.precond = nat.PC[precond]
        sd.mg_levels = int(mg_levels)
        sd.precision = nat.PREC[precision]
        sd.ce_format = nat.CE_FMT[ce_format]
        sd.maxiter_momentum = int(maxiter_momentum)
        sd.maxiter_pressure = int(maxiter_pressure)
        sd.approx_inv_variant = int(approx_inv_variant)
        if gamma_d is not None and len(gamma_d):
            sd.gamma_d, sd.n_gamma_d = arr(gamma_d), len(gamma_d)
        if gamma_p is not None and np.size(gamma_p):
            g = np.atleast_1d(np.asarray(gamma_p, dtype=np.int64))
            sd.gamma_p, sd.n_gamma_p = arr(g), len(g)
        if mg_gamma_p is not None:
            flat = np.concatenate([np.atleast_1d(np.asarray(g, dtype=np.int64)) for g in mg_gamma_p])
            off = np.concatenate([[0], np.cumsum([np.size(g) for g in mg_gamma_p])]).astype(np.int64)
            sd.mg_gamma_p, sd.mg_gamma_p_off = arr(flat), arr(off)
        if bconn is not None and len(bconn):
            sd.bconn, sd.nvec, sd.n_bedge = arr(bconn), arr(nvec), len(bconn)
        if gamma_d_phi is not None and len(gamma_d_phi):
            sd.gamma_d_phi, sd.n_gamma_d_phi = arr(gamma_d_phi), len(gamma_d_phi)
        h = ctypes.c_void_p()
        nat.check(lib.fcfd_ctx_create(ctypes.byref(md), ctypes.byref(sd), nat.stream_handle(), ctypes.byref(h)),
                  "fcfd_ctx_create")
        self._h = h
        self._lib = lib
        self._t = t
        self.n_gamma_d = 0 if gamma_d is None else len(gamma_d)
        lv = ctypes.c_int32()
        nat.check(lib.fcfd_ctx_info(h, None, None, ctypes.byref(lv)), "fcfd_ctx_info")
        self.levels = lv.value

    # ---------------------------------------------------------------------------------- lifetime
    def close(self):
        if getattr(self, "_h", None):
            self._lib.fcfd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sync_stream(self):
        nat.check(self._lib.fcfd_set_stream(self._h, nat.stream_handle()), "fcfd_set_stream")

    def _empty(self, n, dtype=None):
        return self._t.empty(n, dtype=dtype or self._t.float64, device=nat.device())

    # ---------------------------------------------------------------------------------- mesh / constants
    def mesh(self):
        coords = self._empty(2 * self.nn)
        nconn = self._empty(4 * self.ne, self._t.int64)
        self._sync_stream()
        nat.check(self._lib.fcfd_mesh_fill(self._h, nat.ptr(coords), nat.ptr(nconn)), "fcfd_mesh_fill")
        return coords.view(self.nn, 2), nconn.view(self.ne, 4)

    def element_constant(self, which):
        if self.order == 2:
            k = self.nen
            out = (ctypes.c_double * (k * k))()
            nat.check(self._lib.fcfd_element_matrix(
                self._h, {"Me": 0, "Le": 1, "Gx": 2, "Gy": 3, "Gz": 4, "Le_off": 5}[which], out, k * k),
                "fcfd_element_matrix")
            return np.array(out[:]).reshape(k, k)
        if self.ndim == 3:
            out = (ctypes.c_double * 64)()
            nat.check(self._lib.fcfd_element_constant3(
                self._h, {"Me": 0, "Le": 1, "Gx": 2, "Gy": 3, "Gz": 4}[which], out), "fcfd_element_constant3")
            return np.array(out[:]).reshape(8, 8)
        out = (ctypes.c_double * 16)()
        nat.check(self._lib.fcfd_element_constant(self._h, {"Me": 0, "Le": 1, "Gx": 2, "Gy": 3, "Le_off": 4}[which],
                                                  out), "fcfd_element_constant")
        return np.array(out[:]).reshape(4, 4)

    def lumped_mass(self):
        M = self._empty(self.ndim * self.nn)
        self._sync_stream()
        nat.check(self._lib.fcfd_lumped_mass(self._h, nat.ptr(M)), "fcfd_lumped_mass")
        return M

    def laplacian_diag(self, level=0):
        n = ctypes.c_int64()
        nat.check(self._lib.fcfd_level_nodes(self._h, level, ctypes.byref(n), None, None), "fcfd_level_nodes")
        out = self._empty(n.value)
        self._sync_stream()
        nat.check(self._lib.fcfd_laplacian_diag(self._h, level, nat.ptr(out)), "fcfd_laplacian_diag")
        return out

    def level_shape(self, level):
        n, nx, ny = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        nat.check(self._lib.fcfd_level_nodes(self._h, level, ctypes.byref(n), ctypes.byref(nx), ctypes.byref(ny)),
                  "fcfd_level_nodes")
        return n.value, nx.value, ny.value

    # ---------------------------------------------------------------------------------- element ops
    def conv_elem_calc(self, U, beta):
        U = nat.as_dev(U)
        k = self.nen
        Ce = self._empty(k * k * self.ne)
        self._sync_stream()
        nat.check(self._lib.fcfd_conv_elem_calc(self._h, nat.ptr(U), float(beta), nat.ptr(Ce)), "fcfd_conv_elem_calc")
        return Ce.view(self.ne, k, k)

    def mass_matvec(self, x):
        """M_c x per velocity component (consistent_mass_mult, element_mult.py:11-36)."""
        x = nat.as_dev(x)
        y = self._empty(self.ndim * self.nn)
        self._sync_stream()
        nat.check(self._lib.fcfd_mass_matvec(self._h, nat.ptr(x), nat.ptr(y)), "fcfd_mass_matvec")
        return y

    def boundary_integral(self, U_d):
        Fp = self._empty(self.nn)
        self._sync_stream()
        Ud = None if U_d is None else nat.as_dev(U_d)
        nat.check(self._lib.fcfd_boundary_integral(self._h, nat.ptr(Ud), nat.ptr(Fp)), "fcfd_boundary_integral")
        return Fp

    def elem_mult_const(self, A, x, level=0, out=None, accumulate=False):
        x = nat.as_dev(x)
        if out is None:
            out = self._empty(x.numel()) if not accumulate else None
        self._sync_stream()
        nat.check(self._lib.fcfd_elem_mult_const(self._h, int(level), nat.c16(A), nat.ptr(x), nat.ptr(out),
                                                 int(bool(accumulate))), "fcfd_elem_mult_const")
        return out

    def elem_mult_stored(self, x, K0=None, A1=None, c1=0.0, A2=None, c2=0.0, ncomp=1):
        x = nat.as_dev(x)
        A1 = None if A1 is None else nat.as_dev(A1)
        A2 = None if A2 is None else nat.as_dev(A2)
        y = self._empty(self.nn * ncomp)
        self._sync_stream()
        k0 = None if K0 is None else nat.c16(K0)
        nat.check(self._lib.fcfd_elem_mult_stored(self._h, k0, nat.ptr(A1), float(c1), nat.ptr(A2), float(c2),
                                                  int(ncomp), nat.ptr(x), nat.ptr(y)), "fcfd_elem_mult_stored")
        return y

    # ---------------------------------------------------------------------------------- solvers
    def vcycle(self, r):
        r = nat.as_dev(r)
        z = self._empty(self.nn)
        self._sync_stream()
        nat.check(self._lib.fcfd_vcycle(self._h, nat.ptr(r), nat.ptr(z)), "fcfd_vcycle")
        return z

    def level_nn(self, level):
        return self.level_shape(level)[0]

    def interpolate(self, xH, level=0):
        """Prolongation of xH on level + 1 to level (multigrid.py:90-140)."""
        xH = nat.as_dev(xH)
        xh = self._empty(self.level_nn(level))
        self._sync_stream()
        nat.check(self._lib.fcfd_interpolate(self._h, int(level), nat.ptr(xH), nat.ptr(xh)), "fcfd_interpolate")
        return xh

    def restrict(self, xh, level=0):
        """R = I^T of xh on level to level + 1 (multigrid.py:142-212)."""
        xh = nat.as_dev(xh)
        xH = self._empty(self.level_nn(level + 1))
        self._sync_stream()
        nat.check(self._lib.fcfd_restrict(self._h, int(level), nat.ptr(xh), nat.ptr(xH)), "fcfd_restrict")
        return xH

    def jacobi_sweep(self, x, rhs, Le_off, L_diag, gamma_p, set_zeroP, level=0):
        """One damped Jacobi sweep (multigrid.py:214-228) on a level."""
        x, rhs, Ld = nat.as_dev(x), nat.as_dev(rhs), nat.as_dev(L_diag)
        gp = None if gamma_p is None else nat.as_index(gamma_p)
        lo = (ctypes.c_double * 16)(*np.asarray(nat.to_numpy(Le_off), dtype=np.float64).ravel())
        out = self._empty(self.level_nn(level))
        self._sync_stream()
        nat.check(self._lib.fcfd_jacobi_sweep(self._h, int(level), lo, nat.ptr(x), nat.ptr(rhs), nat.ptr(Ld),
                                              nat.ptr(gp), 0 if gp is None else int(gp.numel()),
                                              int(bool(set_zeroP)), nat.ptr(out)), "fcfd_jacobi_sweep")
        return out

    def pressure_cg(self, b, x0=None, tol=None, maxiter=1000):
        b = nat.as_dev(b)
        x = self._t.zeros(self.nn, dtype=self._t.float64, device=nat.device()) if x0 is None else nat.as_dev(x0).clone()
        it, conv = ctypes.c_int32(), ctypes.c_int32()
        self._sync_stream()
        nat.check(self._lib.fcfd_pressure_cg(self._h, nat.ptr(b), nat.ptr(x), float(self.tol if tol is None else tol),
                                             int(maxiter), ctypes.byref(it), ctypes.byref(conv)), "fcfd_pressure_cg")
        return x, it.value, bool(conv.value)

    def solve_pressure(self, U_star, P_n, F_p, rho, dt):
        P = self._empty(self.nn)
        it = ctypes.c_int32()
        U_star, P_n = nat.as_dev(U_star), nat.as_dev(P_n)
        Fp = _opt_vec(F_p, self.nn)
        self._sync_stream()
        nat.check(self._lib.fcfd_solve_pressure(self._h, nat.ptr(U_star), nat.ptr(P_n), nat.ptr(Fp), float(rho),
                                                float(dt), nat.ptr(P), ctypes.byref(it)), "fcfd_solve_pressure")
        return P, it.value

    def solve_momentum(self, U_n, P_n, rho, mu, dt, U_d=None, F=None, Ce=None):
        U_n, P_n = nat.as_dev(U_n), nat.as_dev(P_n)
        Ud = None if U_d is None else nat.as_dev(U_d)
        Fv = _opt_vec(F, self.ndim * self.nn)
        Ce = None if Ce is None else nat.as_dev(Ce)
        out = self._empty(self.ndim * self.nn)
        it = ctypes.c_int32()
        self._sync_stream()
        nat.check(self._lib.fcfd_solve_momentum(self._h, nat.ptr(U_n), nat.ptr(P_n), nat.ptr(Ce), float(rho), float(mu),
                                                float(dt), nat.ptr(Ud), nat.ptr(Fv), nat.ptr(out), ctypes.byref(it)),
                  "fcfd_solve_momentum")
        return out, it.value

    def four_step_timestep(self, U_n, P_n, dt, rho, mu, U_d=None, F_p=None, F=None, out=None):
        U_n, P_n = nat.as_dev(U_n), nat.as_dev(P_n)
        Ud = None if U_d is None else nat.as_dev(U_d)
        Fp = _opt_vec(F_p, self.nn)
        Fv = _opt_vec(F, self.ndim * self.nn)
        if out is None:
            U, P = self._empty(self.ndim * self.nn), self._empty(self.nn)
        else:
            U, P = out
        iters = (ctypes.c_int32 * 2)()
        self._sync_stream()
        nat.check(self._lib.fcfd_four_step_timestep(self._h, nat.ptr(U_n), nat.ptr(P_n), float(dt), float(rho),
                                                    float(mu), nat.ptr(Ud), nat.ptr(Fp), nat.ptr(Fv), nat.ptr(U),
                                                    nat.ptr(P), iters), "fcfd_four_step_timestep")
        return U, P, int(iters[0]), int(iters[1])

    def four_step_rollout(self, U0, P0, nsteps, dt, rho, mu, U_d=None, F_p=None, F=None):
        """nsteps four_step_timestep calls in one native loop (fixed F_p, F, U_d): the compiled
        step loop of timestep_loop_scan (simulation.py:362-374).  Returns (U, P, iters) with
        iters an (nsteps, 2) int array of (U_iters, P_iters)."""
        U0, P0 = nat.as_dev(U0), nat.as_dev(P0)
        Ud = None if U_d is None else nat.as_dev(U_d)
        Fp = _opt_vec(F_p, self.nn)
        Fv = _opt_vec(F, self.ndim * self.nn)
        U, P = self._empty(self.ndim * self.nn), self._empty(self.nn)
        iters = (ctypes.c_int32 * (2 * int(nsteps)))()
        self._sync_stream()
        nat.check(self._lib.fcfd_four_step_rollout(self._h, nat.ptr(U0), nat.ptr(P0), int(nsteps), float(dt),
                                                   float(rho), float(mu), nat.ptr(Ud), nat.ptr(Fp), nat.ptr(Fv),
                                                   nat.ptr(U), nat.ptr(P), iters), "fcfd_four_step_rollout")
        return U, P, np.frombuffer(iters, dtype=np.int32).reshape(int(nsteps), 2).copy()

    def solve_scalar(self, U_n, U, phi_n, dt, D, phi_d=None, supg=False):
        U_n, U, phi_n = nat.as_dev(U_n), nat.as_dev(U), nat.as_dev(phi_n)
        pd = None if phi_d is None else nat.as_dev(phi_d)
        phi = self._empty(self.nn)
        it = ctypes.c_int32()
        self._sync_stream()
        nat.check(self._lib.fcfd_solve_scalar(self._h, nat.ptr(U_n), nat.ptr(U), nat.ptr(phi_n), float(dt), float(D),
                                              nat.ptr(pd), int(bool(supg)), nat.ptr(phi), ctypes.byref(it)),
                  "fcfd_solve_scalar")
        return phi, it.value

    def passive_timestep(self, U_n, P_n, phi_n, dt, rho, mu, D, U_d=None, phi_d=None, F_p=None, F=None,
                         supg=False):
        U_n, P_n, phi_n = nat.as_dev(U_n), nat.as_dev(P_n), nat.as_dev(phi_n)
        Ud = None if U_d is None else nat.as_dev(U_d)
        pd = None if phi_d is None else nat.as_dev(phi_d)
        Fp, Fv = _opt_vec(F_p, self.nn), _opt_vec(F, 2 * self.nn)
        U, P, phi = self._empty(2 * self.nn), self._empty(self.nn), self._empty(self.nn)
        iters = (ctypes.c_int32 * 3)()
        self._sync_stream()
        nat.check(self._lib.fcfd_passive_timestep(self._h, nat.ptr(U_n), nat.ptr(P_n), nat.ptr(phi_n), float(dt),
                                                  float(rho), float(mu), float(D), nat.ptr(Ud), nat.ptr(pd),
                                                  nat.ptr(Fp), nat.ptr(Fv), int(bool(supg)), nat.ptr(U), nat.ptr(P),
                                                  nat.ptr(phi), iters), "fcfd_passive_timestep")
        return U, P, phi, int(iters[0]), int(iters[1]), int(iters[2])

    def four_step_timestep_vjp(self, U_n, P_n, dt, rho, mu, gU=None, gP=None, U_d=None, F_p=None, F=None):
        """Reverse-mode product of four_step_timestep: ((gU_n, gP_n), (U, P), iters[4])."""
        U_n, P_n = nat.as_dev(U_n), nat.as_dev(P_n)
        Ud = None if U_d is None else nat.as_dev(U_d)
        gU = None if gU is None else nat.as_dev(gU)
        gP = None if gP is None else nat.as_dev(gP)
        Fp, Fv = _opt_vec(F_p, self.nn), _opt_vec(F, 2 * self.nn)
        U, P = self._empty(2 * self.nn), self._empty(self.nn)
        gUn, gPn = self._empty(2 * self.nn), self._empty(self.nn)
        iters = (ctypes.c_int32 * 4)()
        self._sync_stream()
        nat.check(self._lib.fcfd_four_step_timestep_vjp(
            self._h, nat.ptr(U_n), nat.ptr(P_n), float(dt), float(rho), float(mu), nat.ptr(Ud), nat.ptr(Fp),
            nat.ptr(Fv), nat.ptr(gU), nat.ptr(gP), nat.ptr(U), nat.ptr(P), nat.ptr(gUn), nat.ptr(gPn), iters),
            "fcfd_four_step_timestep_vjp")
        return (gUn, gPn), (U, P), tuple(int(i) for i in iters)

    def passive_timestep_vjp(self, U_n, P_n, phi_n, dt, rho, mu, D, gU=None, gP=None, gphi=None, U_d=None,
                             phi_d=None, F_p=None, F=None, supg=False, out=None):
        """Reverse-mode product of passive_timestep:
        ((gU_n, gP_n, gphi_n), (U, P, phi), iters[6]).  `out` = optional 6 preallocated
        tensors (U, P, phi, gU_n, gP_n, gphi_n); the cotangent outputs may alias gU, gP, gphi."""
        U_n, P_n, phi_n = nat.as_dev(U_n), nat.as_dev(P_n), nat.as_dev(phi_n)
        Ud = None if U_d is None else nat.as_dev(U_d)
        pd = None if phi_d is None else nat.as_dev(phi_d)
        gU = None if gU is None else nat.as_dev(gU)
        gP = None if gP is None else nat.as_dev(gP)
        gphi = None if gphi is None else nat.as_dev(gphi)
        Fp, Fv = _opt_vec(F_p, self.nn), _opt_vec(F, 2 * self.nn)
        if out is None:
            nn = self.nn
            out = (self._empty(2 * nn), self._empty(nn), self._empty(nn), self._empty(2 * nn), self._empty(nn),
                   self._empty(nn))
        U, P, phi, gUn, gPn, gphin = out
        iters = (ctypes.c_int32 * 6)()
        self._sync_stream()
        nat.check(self._lib.fcfd_passive_timestep_vjp(
            self._h, nat.ptr(U_n), nat.ptr(P_n), nat.ptr(phi_n), float(dt), float(rho), float(mu), float(D),
            nat.ptr(Ud), nat.ptr(pd), nat.ptr(Fp), nat.ptr(Fv), int(bool(supg)), nat.ptr(gU), nat.ptr(gP),
            nat.ptr(gphi), nat.ptr(U), nat.ptr(P), nat.ptr(phi), nat.ptr(gUn), nat.ptr(gPn), nat.ptr(gphin), iters),
            "fcfd_passive_timestep_vjp")
        return (gUn, gPn, gphin), (U, P, phi), tuple(int(i) for i in iters)

    def supg_elem_calc(self, U_n, Ce=None, D=0.0, dt=1.0, r=2.0):
        U_n = nat.as_dev(U_n)
        Ce = None if Ce is None else nat.as_dev(Ce)
        Ct = self._empty(16 * self.ne)
        tau = self._empty(self.ne) if Ce is not None else None
        self._sync_stream()
        nat.check(self._lib.fcfd_supg_elem_calc(self._h, nat.ptr(U_n), nat.ptr(Ce), float(D), float(dt), float(r),
                                                nat.ptr(Ct), nat.ptr(tau)), "fcfd_supg_elem_calc")
        return Ct.view(self.ne, 4, 4), tau

    def dot(self, x, y):
        out = ctypes.c_double()
        self._sync_stream()
        nat.check(self._lib.fcfd_dot(self._h, nat.ptr(x), nat.ptr(y), x.numel(), ctypes.byref(out)), "fcfd_dot")
        return out.value

    def axpby(self, a, x, b, y):
        self._sync_stream()
        nat.check(self._lib.fcfd_axpby(self._h, float(a), nat.ptr(x), float(b), nat.ptr(y), y.numel()), "fcfd_axpby")
        return y

    def bench_momentum_apply(self, reps=20):
        ms = ctypes.c_float()
        self._sync_stream()
        nat.check(self._lib.fcfd_bench_momentum_apply(self._h, int(reps), ctypes.byref(ms)), "fcfd_bench_momentum_apply")
        return ms.value


def _opt_vec(v, n):
    """None for an absent or scalar-zero vector (the reference passes 0)."""
    if v is None:
        return None
    if np.isscalar(v) or (hasattr(v, "ndim") and v.ndim == 0):
        if float(v) == 0.0:
            return None
        return nat.as_dev(np.full(n, float(v)))
    return nat.as_dev(v)


# ------------------------------------------------------------------------------------ grid registry
_grid_cache = {}


def _h_from_shape(wq, Nderiv):
    """Element sizes (hx, hy) of a rectangular quad4 element from its 2x2
    Gauss data (dN1/dx and dN3/dy at the first point)."""
    s = 1.0 / math.sqrt(3.0)
    Nd = nat.to_numpy(Nderiv)
    hx = (1 + s) / (2 * Nd[0, 0, 1])
    hy = (1 + s) / (2 * Nd[0, 1, 3])
    return float(hx), float(hy)


def infer_grid(nconn, nn):
    """(nex, ney, periodic) of the structured Q1 grid behind ``nconn``;
    raises for connectivity that is not one of the reference's regular meshes."""
    nc = nconn
    ne = int(nc.shape[0])
    first = nat.to_numpy(nc[0]).astype(np.int64)
    last = nat.to_numpy(nc[ne - 1]).astype(np.int64)
    if ne == 1:
        raise ValueError("cannot infer a grid from a single element")
    nnx = int(first[3])
    if nn == ne:
        nex, periodic = nnx, True
    else:
        nex, periodic = nnx - 1, False
    if nex < 1 or ne % nex:
        raise ValueError("nconn is not a structured Q1 grid of the reference layout")
    ney = ne // nex
    if (not periodic and nn != (nex + 1) * (ney + 1)) or (periodic and nn != nex * ney):
        raise ValueError("nconn / nn mismatch for a structured Q1 grid")
    ie, je = nex - 1, ney - 1
    if periodic:
        exp = [ie + je * nnx, 0 + je * nnx, 0, ie]
    else:
        n0 = je * nnx + ie
        exp = [n0, n0 + 1, n0 + nnx + 1, n0 + nnx]
    if list(last) != exp or list(first[:3]) != [0, 1, nnx + 1]:
        raise ValueError("nconn is not a structured Q1 grid of the reference layout")
    return nex, ney, periodic


def grid_context(nconn, nn, hx=1.0, hy=1.0):
    """Cached operator-only context for the grid behind nconn."""
    nex, ney, periodic = infer_grid(nconn, nn)
    key = (nex, ney, periodic, round(hx, 15), round(hy, 15))
    ctx = _grid_cache.get(key)
    if ctx is None:
        ctx = Context(nex, ney, hx * nex, hy * ney, periodic, precond=None)
        _grid_cache[key] = ctx
    return ctx
