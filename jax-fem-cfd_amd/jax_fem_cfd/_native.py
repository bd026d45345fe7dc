"""ctypes binding of libfcfd.so (include/fcfd.h) and device-array plumbing.

The HIP library is the only compute path: importing an op without the built
library, or calling one without a GPU, raises.  torch is used for device memory
and the stream handle only; array arguments are float64 / int64 torch CUDA
tensors (NumPy inputs are copied to the device).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FCFD_LIB: another build of the same library (e.g. a variant built elsewhere)
LIB_PATH = os.environ.get("FCFD_LIB") or os.path.join(_HERE, "libfcfd.so")

FCFD_OK = 0
PC = {None: 0, "none": 0, "jacobi": 1, "approx_inv": 2, "multigrid": 3}
PREC = {"fp64": 0, "mixed": 1}
CE_FMT = {"f64": 0, "f32": 1, "f16s": 2, "bf16s": 3}


class NativeError(RuntimeError):
    pass


class MeshDesc(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("nex", ctypes.c_int32), ("ney", ctypes.c_int32),
                ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("periodic", ctypes.c_int32),
                ("nez", ctypes.c_int32), ("lz", ctypes.c_double), ("order", ctypes.c_int32)]


_I64P = ctypes.POINTER(ctypes.c_int64)


class SolverDesc(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("set_zeroP", ctypes.c_int32), ("precond", ctypes.c_int32),
                ("mg_levels", ctypes.c_int32), ("precision", ctypes.c_int32), ("ce_format", ctypes.c_int32),
                ("maxiter_momentum", ctypes.c_int32), ("maxiter_pressure", ctypes.c_int32),
                ("gamma_d", _I64P), ("n_gamma_d", ctypes.c_int64),
                ("gamma_p", _I64P), ("n_gamma_p", ctypes.c_int64),
                ("mg_gamma_p", _I64P), ("mg_gamma_p_off", _I64P),
                ("bconn", _I64P), ("nvec", _I64P), ("n_bedge", ctypes.c_int64),
                ("gamma_d_phi", _I64P), ("n_gamma_d_phi", ctypes.c_int64),
                ("approx_inv_variant", ctypes.c_int32)]


_vp = ctypes.c_void_p
_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64

# name: (restype, argtypes)
_SIGS = {
    "fcfd_ctx_create": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(SolverDesc), _vp, ctypes.POINTER(_vp)]),
    "fcfd_ctx_destroy": (ctypes.c_int, [_vp]),
    "fcfd_set_stream": (ctypes.c_int, [_vp, _vp]),
    "fcfd_last_error": (ctypes.c_char_p, []),
    "fcfd_ctx_info": (ctypes.c_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i32)]),
    "fcfd_level_nodes": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_i64), ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "fcfd_mesh_fill": (ctypes.c_int, [_vp, _vp, _vp]),
    "fcfd_element_constant": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_d)]),
    "fcfd_lumped_mass": (ctypes.c_int, [_vp, _vp]),
    "fcfd_element_constant3": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_d)]),
    "fcfd_mass_matvec": (ctypes.c_int, [_vp, _vp, _vp]),
    "fcfd_element_matrix": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_d), _i32]),
    "fcfd_laplacian_diag": (ctypes.c_int, [_vp, _i32, _vp]),
    "fcfd_conv_elem_calc": (ctypes.c_int, [_vp, _vp, _d, _vp]),
    "fcfd_boundary_integral": (ctypes.c_int, [_vp, _vp, _vp]),
    "fcfd_elem_mult_const": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_d), _vp, _vp, _i32]),
    "fcfd_elem_mult_stored": (ctypes.c_int, [_vp, ctypes.POINTER(_d), _vp, _d, _vp, _d, _i32, _vp, _vp]),
    "fcfd_vcycle": (ctypes.c_int, [_vp, _vp, _vp]),
    "fcfd_interpolate": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "fcfd_restrict": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "fcfd_jacobi_sweep": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_d), _vp, _vp, _vp, _vp, _i64, _i32, _vp]),
    "fcfd_pressure_cg": (ctypes.c_int, [_vp, _vp, _vp, _d, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "fcfd_solve_pressure": (ctypes.c_int, [_vp, _vp, _vp, _vp, _d, _d, _vp, ctypes.POINTER(_i32)]),
    "fcfd_solve_momentum": (ctypes.c_int, [_vp, _vp, _vp, _vp, _d, _d, _d, _vp, _vp, _vp, ctypes.POINTER(_i32)]),
    "fcfd_four_step_timestep": (ctypes.c_int, [_vp, _vp, _vp, _d, _d, _d, _vp, _vp, _vp, _vp, _vp,
                                               ctypes.POINTER(_i32)]),
    "fcfd_four_step_rollout": (ctypes.c_int, [_vp, _vp, _vp, _i32, _d, _d, _d, _vp, _vp, _vp, _vp, _vp,
                                              ctypes.POINTER(ctypes.c_int32)]),
    "fcfd_solve_scalar": (ctypes.c_int, [_vp, _vp, _vp, _vp, _d, _d, _vp, _i32, _vp, ctypes.POINTER(_i32)]),
    "fcfd_passive_timestep": (ctypes.c_int, [_vp, _vp, _vp, _vp, _d, _d, _d, _d, _vp, _vp, _vp, _vp, _i32, _vp, _vp,
                                             _vp, ctypes.POINTER(_i32)]),
    "fcfd_supg_elem_calc": (ctypes.c_int, [_vp, _vp, _vp, _d, _d, _d, _vp, _vp]),
    "fcfd_four_step_timestep_vjp": (ctypes.c_int, [_vp, _vp, _vp, _d, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                   _vp, _vp, ctypes.POINTER(_i32)]),
    "fcfd_passive_timestep_vjp": (ctypes.c_int, [_vp, _vp, _vp, _vp, _d, _d, _d, _d, _vp, _vp, _vp, _vp, _i32, _vp,
                                                 _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_i32)]),
    "fcfd_dot": (ctypes.c_int, [_vp, _vp, _vp, _i64, ctypes.POINTER(_d)]),
    "fcfd_dot_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp]),
    "fcfd_axpby": (ctypes.c_int, [_vp, _d, _vp, _d, _vp, _i64]),
    "fcfd_bench_momentum_apply": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load libfcfd.so and declare every entry point of include/fcfd.h.  Works
    without a GPU (symbols only)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise NativeError(f"libfcfd.so not built at {path}: run __graft_entry__.build() "
                          "(or make -C jax-fem-cfd_amd)")
    # torch ships its own HIP runtime (SONAME libamdhip64.so.7).  Loading it first makes
    # libfcfd bind to that same runtime instead of pulling a second copy from /opt/rocm
    # into the process (two HIP runtimes in one process do not share the device).
    import torch as _t  # noqa: F401
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)


def check(rc, what=""):
    if rc != FCFD_OK:
        msg = load_library().fcfd_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed ({rc}): {msg}")


# ------------------------------------------------------------------------------------ torch plumbing
def torch():
    import torch as _t
    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise NativeError("the HIP product path needs a GPU (torch.cuda.is_available() is False)")
    return t


def device():
    return require_gpu().device("cuda", 0)


def stream_handle():
    t = require_gpu()
    return ctypes.c_void_p(t.cuda.current_stream().cuda_stream)


def as_dev(x, dtype=None):
    """float64 (or given dtype) contiguous CUDA tensor from a tensor / array / scalar."""
    t = require_gpu()
    if dtype is None:
        dtype = t.float64
    if isinstance(x, t.Tensor):
        y = x.to(device=device(), dtype=dtype)
        return y.contiguous()
    return t.as_tensor(np.asarray(x), dtype=dtype, device=device()).contiguous()


def as_index(x):
    """int64 contiguous CUDA tensor of node / dof indices."""
    t = require_gpu()
    return as_dev(np.asarray(to_numpy(x), dtype=np.int64), dtype=t.int64)


def ptr(x):
    return ctypes.c_void_p(0 if x is None else x.data_ptr())


def to_numpy(x):
    t = torch()
    if isinstance(x, t.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def i64_array(a):
    """(ctypes pointer, keepalive ndarray) for an int64 host array or None."""
    if a is None:
        return None, None
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.int64).ravel())
    return arr.ctypes.data_as(_I64P), arr


def c16(A):
    arr = np.ascontiguousarray(np.asarray(to_numpy(A), dtype=np.float64).reshape(16))
    return (_d * 16)(*arr.tolist())
