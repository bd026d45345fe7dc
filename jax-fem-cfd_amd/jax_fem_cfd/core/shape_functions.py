"""Gauss-point shape-function data ``(wq, N, Nderiv)`` (host constants).

Same functions as jax_fem_cfd/core/shape_functions.py.  They are evaluated once,
on element 0, when the solver is set up (standard.py:211); the GPU kernels take
the resulting 4x4 element constants.  I = node, q = quadrature point,
d = dimension; N (q, I), Nderiv (q, d, I), wq (q, 1).
"""
import numpy as np

_S3 = 1.0 / np.sqrt(3.0)


def lin2_shape_functions(coords):
    """shape_functions.py:11-25"""
    coords = np.asarray(coords, dtype=float)
    xi = np.array([-1.0, 1.0]) * _S3
    N = np.stack([(1 - xi) / 2, (1 + xi) / 2], axis=1)
    dNdxi = np.array([[-0.5, 0.5], [-0.5, 0.5]])
    d = coords[1] - coords[0]
    h = np.sqrt(d[0] ** 2 + d[1] ** 2)
    return (np.ones(2) * (h / 2))[:, None], N, (dNdxi * (2 / h))[:, None, :]


def quad4_shape_functions(coords):
    """shape_functions.py:27-45: 2x2 Gauss, q-order (-,-),(+,-),(+,+),(-,+), wq = det J."""
    coords = np.asarray(coords, dtype=float)
    xi = np.array([-1.0, 1.0, 1.0, -1.0]) * _S3
    eta = np.array([-1.0, -1.0, 1.0, 1.0]) * _S3
    N = np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)],
                 axis=1) / 4
    dNdxi = np.stack([-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)], axis=1) / 4
    dNdeta = np.stack([-(1 - xi), -(1 + xi), 1 + xi, 1 - xi], axis=1) / 4
    J = np.stack([dNdxi @ coords, dNdeta @ coords], axis=1)
    return np.linalg.det(J)[:, None], N, np.linalg.inv(J) @ np.stack([dNdxi, dNdeta], axis=1)


def hex8_shape_functions(coords):
    """shape_functions.py:47-86 (host data only; the 3D solver is not built)."""
    coords = np.asarray(coords, dtype=float)
    sx = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float)
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float)
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float)
    xi, eta, zeta = sx * _S3, sy * _S3, sz * _S3
    fx, fy, fz = 1 + xi[:, None] * sx, 1 + eta[:, None] * sy, 1 + zeta[:, None] * sz
    N = fx * fy * fz / 8
    dN = np.stack([sx * fy * fz / 8, fx * sy * fz / 8, fx * fy * sz / 8], axis=1)
    J = np.stack([dN[:, 0] @ coords, dN[:, 1] @ coords, dN[:, 2] @ coords], axis=1)
    return np.linalg.det(J)[:, None], N, np.linalg.inv(J) @ dN


def lin3_shape_functions(coords):
    """shape_functions.py:89-106 (host data only)."""
    coords = np.asarray(coords, dtype=float)
    xi = np.array([-np.sqrt(3 / 5), 0.0, np.sqrt(3 / 5)])
    w = np.array([5 / 9, 8 / 9, 5 / 9])
    N = np.stack([xi * (xi - 1) / 2, 1 - xi ** 2, xi * (xi + 1) / 2], axis=1)
    dNdxi = np.stack([(2 * xi - 1) / 2, -2 * xi, (2 * xi + 1) / 2], axis=1)
    J = np.sqrt((dNdxi @ coords[:, 0]) ** 2 + (dNdxi @ coords[:, 1]) ** 2)
    return (w * J)[:, None], N, (dNdxi / J[:, None])[:, None, :]


def quad9_shape_functions(coords):
    """shape_functions.py:108-159 (host data only; the Q2 solver is not built)."""
    coords = np.asarray(coords, dtype=float)
    g = np.array([-np.sqrt(3 / 5), 0.0, np.sqrt(3 / 5)])
    xi, eta = (a.ravel() for a in np.meshgrid(g, g))
    w1 = np.array([5 / 9, 8 / 9, 5 / 9])
    w = np.outer(w1, w1).ravel()
    N = np.stack([xi * eta * (xi - 1) * (eta - 1) / 4, (1 - xi ** 2) * eta * (eta - 1) / 2,
                  xi * eta * (xi + 1) * (eta - 1) / 4, xi * (xi + 1) * (1 - eta ** 2) / 2,
                  xi * eta * (xi + 1) * (eta + 1) / 4, (1 - xi ** 2) * eta * (eta + 1) / 2,
                  xi * eta * (xi - 1) * (eta + 1) / 4, xi * (xi - 1) * (1 - eta ** 2) / 2,
                  (1 - xi ** 2) * (1 - eta ** 2)], axis=1)
    dNdxi = np.stack([eta * (2 * xi - 1) * (eta - 1) / 4, -2 * xi * eta * (eta - 1) / 2,
                      eta * (2 * xi + 1) * (eta - 1) / 4, (2 * xi + 1) * (1 - eta ** 2) / 2,
                      eta * (2 * xi + 1) * (eta + 1) / 4, -2 * xi * eta * (eta + 1) / 2,
                      eta * (2 * xi - 1) * (eta + 1) / 4, (2 * xi - 1) * (1 - eta ** 2) / 2,
                      -2 * xi * (1 - eta ** 2)], axis=1)
    dNdeta = np.stack([xi * (xi - 1) * (2 * eta - 1) / 4, (1 - xi ** 2) * (2 * eta - 1) / 2,
                       xi * (xi + 1) * (2 * eta - 1) / 4, xi * (xi + 1) * (-2 * eta) / 2,
                       xi * (xi + 1) * (2 * eta + 1) / 4, (1 - xi ** 2) * (2 * eta + 1) / 2,
                       xi * (xi - 1) * (2 * eta + 1) / 4, xi * (xi - 1) * (-2 * eta) / 2,
                       -2 * eta * (1 - xi ** 2)], axis=1)
    J = np.stack([dNdxi @ coords, dNdeta @ coords], axis=1)
    return (w * np.linalg.det(J))[:, None], N, np.linalg.inv(J) @ np.stack([dNdxi, dNdeta], axis=1)
