"""fcfd_dot_noalloc (the allocation-free variant of fcfd_dot) through the C ABI.

Its value must equal fcfd_dot's bitwise (same kernels, same fixed summation order), over two
calls in a row on one context, the second with a smaller n so that stale partials of the first
call would show; n beyond FCFD_DOT_NOALLOC_MAX_N is rejected on the host before any GPU work.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_N = 1 << 24  # FCFD_DOT_NOALLOC_MAX_N in include/fcfd.h


def test_dot_noalloc_matches_dot_on_two_calls():
    from jax_fem_cfd import _native as nat
    from jax_fem_cfd._context import Context
    t = nat.torch()
    lib = nat.load_library()
    ctx = Context(8, 8, 1.0, 1.0, False, precond=None)
    rng = np.random.default_rng(11)
    out = t.full((2,), np.nan, dtype=t.float64, device=nat.device())
    for n in (2 * 1025 * 1025, 4099):
        print(f"\n[case] fcfd_dot_noalloc n={n}", flush=True)
        x = nat.as_dev(rng.standard_normal(n))
        y = nat.as_dev(rng.standard_normal(n))
        ref = ctypes.c_double()
        ctx._sync_stream()
        nat.check(lib.fcfd_dot(ctx._h, nat.ptr(x), nat.ptr(y), n, ctypes.byref(ref)), "fcfd_dot")
        nat.check(lib.fcfd_dot_noalloc(ctx._h, nat.ptr(x), nat.ptr(y), n, ctypes.c_void_p(out.data_ptr())),
                  "fcfd_dot_noalloc")
        got = out[0].item()  # torch's stream; the library's work is ordered before it on return
        assert got == ref.value, (got, ref.value)
        assert abs(got - float(np.dot(x.cpu().numpy(), y.cpu().numpy()))) <= 1e-12 * n
        assert np.isnan(out[1].item())  # nothing beyond the one output value is written


def test_dot_noalloc_rejects_oversize_and_bad_arguments():
    from jax_fem_cfd import _native as nat
    from jax_fem_cfd._context import Context
    t = nat.torch()
    lib = nat.load_library()
    ctx = Context(8, 8, 1.0, 1.0, False, precond=None)
    x = t.zeros(4, dtype=t.float64, device=nat.device())
    out = t.zeros(1, dtype=t.float64, device=nat.device())
    print("\n[case] fcfd_dot_noalloc argument checks (host side, no kernel launched)", flush=True)
    rc = lib.fcfd_dot_noalloc(ctx._h, nat.ptr(x), nat.ptr(x), MAX_N + 1, nat.ptr(out))
    assert rc == -1 and b"exceeds FCFD_DOT_NOALLOC_MAX_N" in lib.fcfd_last_error()
    assert lib.fcfd_dot_noalloc(ctx._h, nat.ptr(x), nat.ptr(x), -1, nat.ptr(out)) == -1
    assert lib.fcfd_dot_noalloc(ctx._h, None, nat.ptr(x), 4, nat.ptr(out)) == -1
    assert lib.fcfd_dot_noalloc(ctx._h, nat.ptr(x), nat.ptr(x), 4, None) == -1
