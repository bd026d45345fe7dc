"""Solution and iteration-history files, same formats as
jax_fem_cfd/utils/running.py:16-38: ``.npy`` holding ``[len(U), U..., p...]``
and a (3, n) stack of momentum iterations, pressure iterations and steps."""
import os

import numpy as np

from .. import _native as nat


def use_single_gpu(DEVICE_ID):
    """running.py:4-10: restrict the visible devices (HIP_VISIBLE_DEVICES on ROCm).
    Call before the first GPU use."""
    os.environ["HIP_VISIBLE_DEVICES"] = "" if DEVICE_ID is None else str(DEVICE_ID)


def print_device_info():
    t = nat.torch()
    n = t.cuda.device_count()
    print(f"{n} HIP device(s):", ", ".join(t.cuda.get_device_name(i) for i in range(n)))


def save_sol(U, p, sol_path):
    U, p = nat.to_numpy(U), nat.to_numpy(p)
    sol = np.concatenate((np.array([U.shape[0]]), np.asarray(U), np.asarray(p)))
    np.save(sol_path, sol)
    print(f"solution (U: {U.shape[0]} values, p: {p.shape[0]}) written to {sol_path}")


def load_sol(sol_path):
    sol = np.load(sol_path)
    U = sol[1:(1 + int(sol[0]))]
    p = sol[(1 + int(sol[0])):]
    print(f"solution read from {sol_path}")
    return U, p


def save_iter_hist(momentum_list, pressure_list, steps, path):
    hist = np.stack((momentum_list, pressure_list, steps), axis=0)
    np.save(path, hist)
    print(f"iteration history ({hist.shape[1]} saves) written to {path}")


def load_iter_hist(path):
    hist = np.load(path)
    print(f"iteration history read from {path}")
    return hist[0, :], hist[1, :], hist[2, :]
