"""Structured meshes with the reference's numbering (host index arrays).

Same functions and return tuple as jax_fem_cfd/setup/regular_mesh.py:
``(node_coords, nconn, surfnodes, bconns, nvecs)``.  These are set-up index
arrays (NumPy, int64 / float64); the GPU kernels never read them -- they use
the same index math in closed form, and ``Context.mesh()`` returns the
device-side coordinates/connectivity for a bit-exact comparison.
"""
import numpy as np


def create_mesh_2d(num_elem_x, num_elem_y, domain_size_x, domain_size_y):
    """regular_mesh.py:5-43: node j*nnx + i at (i dx, j dy); element
    je*nex + ie with CCW nodes; faces [down, right, top, left]."""
    nex, ney = int(num_elem_x), int(num_elem_y)
    nnx, nny = nex + 1, ney + 1
    dx, dy = domain_size_x / nex, domain_size_y / ney
    jj, ii = np.meshgrid(np.arange(nny), np.arange(nnx), indexing="ij")
    node_coords = np.stack([ii * dx, jj * dy], axis=-1).reshape(-1, 2)
    je, ie = np.meshgrid(np.arange(ney), np.arange(nex), indexing="ij")
    n0 = je * nnx + ie
    nconn = np.stack([n0, n0 + 1, n0 + nnx + 1, n0 + nnx], axis=-1).reshape(-1, 4).astype(np.int64)
    surfnodes = [np.arange(nnx, dtype=np.int64),
                 (nnx * np.arange(1, nny + 1) - 1).astype(np.int64),
                 (nnx * (nny - 1) + np.arange(nnx)).astype(np.int64),
                 (nnx * np.arange(nny)).astype(np.int64)]
    bconns = [np.stack([s[:-1], s[1:]], axis=1) for s in surfnodes]
    normals = np.array([[0, -1], [1, 0], [0, 1], [-1, 0]], dtype=np.int64)
    nvecs = [np.tile(normals[i], (bconns[i].shape[0], 1)) for i in range(4)]
    return node_coords, nconn, surfnodes, bconns, nvecs


def create_periodic_mesh_2d(nex, ney, Lx, Ly):
    """regular_mesh.py:149-180: nnx = nex, nny = ney, wrap-around connectivity."""
    nex, ney = int(nex), int(ney)
    nnx, nny = nex, ney
    dx, dy = Lx / nex, Ly / ney
    jj, ii = np.meshgrid(np.arange(nny), np.arange(nnx), indexing="ij")
    node_coords = np.stack([ii * dx, jj * dy], axis=-1).reshape(-1, 2)
    e = np.arange(nex * ney)
    ie, je = e % nex, e // nex
    nconn = np.stack([ie + je * nnx, (ie + 1) % nnx + je * nnx, (ie + 1) % nnx + ((je + 1) % nny) * nnx,
                      ie + ((je + 1) % nny) * nnx], axis=1).astype(np.int64)
    return node_coords, nconn, None, None, None


def create_mesh_3d(ne_x, ne_y, ne_z, domain_size_x, domain_size_y, domain_size_z):
    """regular_mesh.py:45-105 (node k*nnx*nny + j*nnx + i)."""
    nnx, nny, nnz = ne_x + 1, ne_y + 1, ne_z + 1
    dx, dy, dz = domain_size_x / ne_x, domain_size_y / ne_y, domain_size_z / ne_z
    k, j, i = np.meshgrid(np.arange(nnz), np.arange(nny), np.arange(nnx), indexing="ij")
    node_coords = np.stack([i * dx, j * dy, k * dz], axis=-1).reshape(-1, 3)
    ke, je, ie = np.meshgrid(np.arange(ne_z), np.arange(ne_y), np.arange(ne_x), indexing="ij")
    q = je * nnx + ie
    quad = np.stack([q, q + 1, q + nnx + 1, q + nnx], axis=-1)
    lay = nnx * nny
    nconn = np.concatenate([ke[..., None] * lay + quad, (ke[..., None] + 1) * lay + quad],
                           axis=-1).reshape(-1, 8).astype(np.int64)
    n3 = nconn.reshape(ne_z, ne_y, ne_x, 8)
    # mixed basic/advanced indexing as regular_mesh.py:77-82 (NumPy semantics)
    bconns = [n3[:, 0, :, [0, 1, 5, 4]].reshape(-1, 4), n3[:, :, -1, [1, 5, 6, 2]].reshape(-1, 4),
              n3[:, -1, :, [3, 2, 6, 7]].reshape(-1, 4), n3[:, :, 0, [0, 4, 7, 3]].reshape(-1, 4),
              n3[0, :, :, [0, 1, 2, 3]].reshape(-1, 4), n3[-1, :, :, [4, 5, 6, 7]].reshape(-1, 4)]
    zx = np.repeat(np.arange(nnz) * lay, nnx)
    zy = np.repeat(np.arange(nnz) * lay, nny)
    surfnodes = [np.tile(np.arange(nnx), nnz) + zx, np.tile(nnx * np.arange(1, nny + 1) - 1, nnz) + zy,
                 np.tile(nnx * (nny - 1) + np.arange(nnx), nnz) + zx, np.tile(nnx * np.arange(nny), nnz) + zy,
                 np.arange(lay), lay * (nnz - 1) + np.arange(lay)]
    surfnodes = [s.astype(np.int64) for s in surfnodes]
    normals = np.array([[0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, -1], [0, 0, 1]])
    nvecs = [np.tile(normals[f], (bconns[f].shape[0], 1)) for f in range(6)]
    return node_coords, nconn, surfnodes, bconns, nvecs


def create_mesh_3d_hex27(ne_x, ne_y, ne_z, domain_size_x, domain_size_y, domain_size_z):
    """Triquadratic 27-node box mesh (the reference has none: simulation.py:73-74).  Node grid
    (2 ne + 1) per direction at half spacing, node k*nnx*nny + j*nnx + i; element
    (ie, je, ke) -> (ke*ne_y + je)*ne_x + ie with local node (c*3 + b)*3 + a at (2ie + a, 2je + b,
    2ke + c).  Faces [D, R, T, L, F, B] list their nodes as create_mesh_3d does; each face element
    is a quad9 patch (corners counterclockwise, mid-sides, centre) over (s, t) = (z, y) on x-faces,
    (x, z) on y-faces, (x, y) on z-faces, first in-face element index fastest."""
    nnx, nny, nnz = 2 * ne_x + 1, 2 * ne_y + 1, 2 * ne_z + 1
    hx, hy, hz = domain_size_x / ne_x / 2, domain_size_y / ne_y / 2, domain_size_z / ne_z / 2
    k, j, i = np.meshgrid(np.arange(nnz), np.arange(nny), np.arange(nnx), indexing="ij")
    node_coords = np.stack([i * hx, j * hy, k * hz], axis=-1).reshape(-1, 3)
    lay = nnx * nny
    ke, je, ie = np.meshgrid(np.arange(ne_z), np.arange(ne_y), np.arange(ne_x), indexing="ij")
    corner = (2 * ke * lay + 2 * je * nnx + 2 * ie).reshape(-1, 1)
    loc = np.array([c * lay + b * nnx + a for c in range(3) for b in range(3) for a in range(3)])
    nconn = (corner + loc[None, :]).astype(np.int64)
    s9 = np.array([0, 1, 2, 2, 2, 1, 0, 0, 1])
    t9 = np.array([0, 0, 0, 1, 2, 2, 2, 1, 1])

    def patches(n1, n2, node):
        e2, e1 = np.meshgrid(np.arange(n2), np.arange(n1), indexing="ij")
        e1, e2 = e1.reshape(-1, 1), e2.reshape(-1, 1)
        return node(e1, e2).astype(np.int64)
    at = lambda x, y, z: z * lay + y * nnx + x
    bconns = [patches(ne_x, ne_z, lambda a, c: at(2 * a + s9, 0, 2 * c + t9)),
              patches(ne_y, ne_z, lambda b, c: at(nnx - 1, 2 * b + t9, 2 * c + s9)),
              patches(ne_x, ne_z, lambda a, c: at(2 * a + s9, nny - 1, 2 * c + t9)),
              patches(ne_y, ne_z, lambda b, c: at(0, 2 * b + t9, 2 * c + s9)),
              patches(ne_x, ne_y, lambda a, b: at(2 * a + s9, 2 * b + t9, 0)),
              patches(ne_x, ne_y, lambda a, b: at(2 * a + s9, 2 * b + t9, nnz - 1))]
    zx = np.repeat(np.arange(nnz) * lay, nnx)
    zy = np.repeat(np.arange(nnz) * lay, nny)
    surfnodes = [np.tile(np.arange(nnx), nnz) + zx, np.tile(nnx * np.arange(1, nny + 1) - 1, nnz) + zy,
                 np.tile(nnx * (nny - 1) + np.arange(nnx), nnz) + zx, np.tile(nnx * np.arange(nny), nnz) + zy,
                 np.arange(lay), lay * (nnz - 1) + np.arange(lay)]
    surfnodes = [x.astype(np.int64) for x in surfnodes]
    normals = np.array([[0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, -1], [0, 0, 1]])
    nvecs = [np.tile(normals[f], (bconns[f].shape[0], 1)) for f in range(6)]
    return node_coords, nconn, surfnodes, bconns, nvecs


def create_mesh_2d_quad9(nex, ney, Lx, Ly):
    """regular_mesh.py:107-145 (index arrays only; the Q2 solver is not built)."""
    nnx, nny = 2 * nex + 1, 2 * ney + 1
    dx, dy = Lx / nex, Ly / ney
    jj, ii = np.meshgrid(np.arange(nny), np.arange(nnx), indexing="ij")
    node_coords = np.stack([ii * dx / 2, jj * dy / 2], axis=-1).reshape(-1, 2)
    e = np.arange(nex * ney)
    ie, je = e % nex, e // nex
    b = 2 * ie + 2 * je * nnx
    nconn = np.stack([b, b + 1, b + 2, b + 2 + nnx, b + 2 + 2 * nnx, b + 1 + 2 * nnx, b + 2 * nnx, b + nnx,
                      b + 1 + nnx], axis=1).astype(np.int64)
    surfnodes = [np.arange(0, nnx), np.arange(nnx - 1, nnx * nny, nnx), np.arange(nnx * (nny - 1), nnx * nny),
                 np.arange(0, nnx * nny, nnx)]
    bconns = [np.stack([s[:-2:2], s[1:-1:2], s[2::2]], axis=1) for s in surfnodes]
    normals = np.array([[0, -1], [1, 0], [0, 1], [-1, 0]])
    nvecs = [np.tile(normals[i], (bconns[i].shape[0], 1)) for i in range(4)]
    return node_coords, nconn, surfnodes, bconns, nvecs


def create_periodic_mesh_2d_quad9(nex, ney, Lx, Ly):
    """regular_mesh.py:182-218 (index arrays only)."""
    nnx, nny = 2 * nex, 2 * ney
    dx, dy = Lx / nex, Ly / ney
    jj, ii = np.meshgrid(np.arange(nny), np.arange(nnx), indexing="ij")
    node_coords = np.stack([ii * dx / 2, jj * dy / 2], axis=-1).reshape(-1, 2)
    e = np.arange(nex * ney)
    ie, je = e % nex, e // nex
    c0, c1, c2 = 2 * ie, 2 * ie + 1, (2 * ie + 2) % nnx
    r0, r1, r2 = 2 * je * nnx, (2 * je + 1) * nnx, ((2 * je + 2) % nny) * nnx
    nconn = np.stack([c0 + r0, c1 + r0, c2 + r0, c2 + r1, c2 + r2, c1 + r2, c0 + r2, c0 + r1, c1 + r1],
                     axis=1).astype(np.int64)
    return node_coords, nconn, None, None, None
