"""Structured Q1 / Q2 meshes and boundary-condition index sets (oracle, test-only).

Follows jax_fem_cfd/setup/regular_mesh.py and jax_fem_cfd/setup/simulation.py.
"""
import numpy as np


def create_mesh_2d(nex, ney, lx, ly):
    """Rectangular Q1 mesh.  regular_mesh.py:5-43.

    Node (i, j) -> j*nnx + i at (i*dx, j*dy); element (ie, je) -> je*nex + ie
    with counter-clockwise nodes [BL, BR, TR, TL].  Faces [down, right, top,
    left]; bconns are consecutive node pairs along each face; outward normals.
    """
    nnx, nny = nex + 1, ney + 1
    dx, dy = lx / nex, ly / ney
    i = np.arange(nnx)
    j = np.arange(nny)
    coords = np.stack(np.broadcast_arrays(i[None, :] * dx, j[:, None] * dy), axis=-1).reshape(-1, 2)
    ie = np.arange(nex)[None, :]
    je = np.arange(ney)[:, None]
    n0 = je * nnx + ie
    nconn = np.stack([n0, n0 + 1, n0 + nnx + 1, n0 + nnx], axis=-1).reshape(-1, 4).astype(np.int64)
    surf = [np.arange(nnx, dtype=np.int64),
            (nnx * np.arange(1, nny + 1) - 1).astype(np.int64),
            (nnx * (nny - 1) + np.arange(nnx)).astype(np.int64),
            (nnx * np.arange(nny)).astype(np.int64)]
    bconns = [np.stack([s[:-1], s[1:]], axis=1) for s in surf]
    normals = np.array([[0, -1], [1, 0], [0, 1], [-1, 0]], dtype=np.int64)
    nvecs = [np.tile(normals[f], (bconns[f].shape[0], 1)) for f in range(4)]
    return coords, nconn, surf, bconns, nvecs


def create_periodic_mesh_2d(nex, ney, lx, ly):
    """Doubly periodic Q1 mesh, no duplicated nodes.  regular_mesh.py:149-180."""
    nnx, nny = nex, ney
    dx, dy = lx / nex, ly / ney
    i = np.arange(nnx)
    j = np.arange(nny)
    coords = np.stack(np.broadcast_arrays(i[None, :] * dx, j[:, None] * dy), axis=-1).reshape(-1, 2)
    e = np.arange(nex * ney)
    ie, je = e % nex, e // nex
    n0 = ie + je * nnx
    n1 = (ie + 1) % nnx + je * nnx
    n2 = (ie + 1) % nnx + ((je + 1) % nny) * nnx
    n3 = ie + ((je + 1) % nny) * nnx
    nconn = np.stack([n0, n1, n2, n3], axis=1).astype(np.int64)
    return coords, nconn, None, None, None


def create_mesh_3d(nex, ney, nez, lx, ly, lz):
    """Box Q1 hex mesh.  regular_mesh.py:45-105 (node k*nnx*nny + j*nnx + i;
    hex nodes = quad at k ++ quad at k+1; faces [D, R, T, L, F, B])."""
    nnx, nny, nnz = nex + 1, ney + 1, nez + 1
    dx, dy, dz = lx / nex, ly / ney, lz / nez
    k, j, i = np.meshgrid(np.arange(nnz), np.arange(nny), np.arange(nnx), indexing="ij")
    coords = np.stack([i * dx, j * dy, k * dz], axis=-1).reshape(-1, 3)
    ke, je, ie = np.meshgrid(np.arange(nez), np.arange(ney), np.arange(nex), indexing="ij")
    q = je * nnx + ie
    quad = np.stack([q, q + 1, q + nnx + 1, q + nnx], axis=-1)
    lay = nnx * nny
    nconn = np.concatenate([ke[..., None] * lay + quad, (ke[..., None] + 1) * lay + quad],
                           axis=-1).reshape(-1, 8).astype(np.int64)
    n3 = nconn.reshape(nez, ney, nex, 8)
    # Same mixed basic/advanced indexing as regular_mesh.py:77-82.  Under
    # NumPy (and JAX) rules the face-local axis moves to the FRONT when the
    # two advanced indices are separated by a slice (faces D, T, F, B), so the
    # reshaped rows of those faces are not per-face-element node quadruples.
    # The quirk is restated as is.
    bD = n3[:, 0, :, [0, 1, 5, 4]].reshape(-1, 4)
    bR = n3[:, :, -1, [1, 5, 6, 2]].reshape(-1, 4)
    bT = n3[:, -1, :, [3, 2, 6, 7]].reshape(-1, 4)
    bL = n3[:, :, 0, [0, 4, 7, 3]].reshape(-1, 4)
    bF = n3[0, :, :, [0, 1, 2, 3]].reshape(-1, 4)
    bB = n3[-1, :, :, [4, 5, 6, 7]].reshape(-1, 4)
    bconns = [bD, bR, bT, bL, bF, bB]
    zoff = np.repeat(np.arange(nnz) * lay, nnx)
    zoffy = np.repeat(np.arange(nnz) * lay, nny)
    surf = [np.tile(np.arange(nnx), nnz) + zoff,
            np.tile(nnx * np.arange(1, nny + 1) - 1, nnz) + zoffy,
            np.tile(nnx * (nny - 1) + np.arange(nnx), nnz) + zoff,
            np.tile(nnx * np.arange(nny), nnz) + zoffy,
            np.arange(lay),
            lay * (nnz - 1) + np.arange(lay)]
    surf = [s.astype(np.int64) for s in surf]
    normals = np.array([[0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, -1], [0, 0, 1]])
    nvecs = [np.tile(normals[f], (bconns[f].shape[0], 1)) for f in range(6)]
    return coords, nconn, surf, bconns, nvecs


# quad9 local node order over face / element coordinates (s, t), as create_mesh_2d_quad9
Q9_POS = [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (0, 1), (1, 1)]


def create_mesh_3d_hex27(nex, ney, nez, lx, ly, lz):
    """Triquadratic 27-node box mesh.  The reference has no 3D quadratic mesh
    (simulation.py:73-74 rejects shape_func_ord = 2 in 3D); this one extends create_mesh_3d
    (regular_mesh.py:45-105) the way create_mesh_2d_quad9 extends create_mesh_2d:
    node grid (2 nex + 1) x (2 ney + 1) x (2 nez + 1) at half spacing, node k*nnx*nny + j*nnx + i
    at (i dx/2, j dy/2, k dz/2); element (ie, je, ke) -> (ke*ney + je)*nex + ie owns the nodes
    (2 ie + a, 2 je + b, 2 ke + c) with local index (c*3 + b)*3 + a.  Faces [D, R, T, L, F, B]
    (y = 0, x = lx, y = ly, x = 0, z = 0, z = lz) list their nodes in create_mesh_3d's order; a
    face element is a quad9 patch in quad9 local order over (s, t) = (z, y) on x-faces (counter-
    clockwise in the (z, y) coordinates the face integral maps from), (x, z) on y-faces and (x, y)
    on z-faces, ordered with the first in-face element index fastest."""
    nnx, nny, nnz = 2 * nex + 1, 2 * ney + 1, 2 * nez + 1
    hx, hy, hz = lx / nex / 2, ly / ney / 2, lz / nez / 2
    k, j, i = np.meshgrid(np.arange(nnz), np.arange(nny), np.arange(nnx), indexing="ij")
    coords = np.stack([i * hx, j * hy, k * hz], axis=-1).reshape(-1, 3)
    lay = nnx * nny
    ke, je, ie = np.meshgrid(np.arange(nez), np.arange(ney), np.arange(nex), indexing="ij")
    base = (2 * ke * lay + 2 * je * nnx + 2 * ie).reshape(-1, 1)
    c, b, a = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    nconn = (base + (c * lay + b * nnx + a).ravel()[None, :]).astype(np.int64)
    node = lambda x, y, z: z * lay + y * nnx + x

    def faces(n1, n2, f):
        rows = []
        for t2 in range(n2):
            for t1 in range(n1):
                rows.append([f(t1, t2, ps, pt) for ps, pt in Q9_POS])
        return np.array(rows, dtype=np.int64)
    # x-faces: s = z, t = y (elements ordered je fastest, then ke)
    bR = faces(ney, nez, lambda je_, ke_, ps, pt: node(nnx - 1, 2 * je_ + pt, 2 * ke_ + ps))
    bL = faces(ney, nez, lambda je_, ke_, ps, pt: node(0, 2 * je_ + pt, 2 * ke_ + ps))
    # y-faces: s = x, t = z
    bD = faces(nex, nez, lambda ie_, ke_, ps, pt: node(2 * ie_ + ps, 0, 2 * ke_ + pt))
    bT = faces(nex, nez, lambda ie_, ke_, ps, pt: node(2 * ie_ + ps, nny - 1, 2 * ke_ + pt))
    # z-faces: s = x, t = y
    bF = faces(nex, ney, lambda ie_, je_, ps, pt: node(2 * ie_ + ps, 2 * je_ + pt, 0))
    bB = faces(nex, ney, lambda ie_, je_, ps, pt: node(2 * ie_ + ps, 2 * je_ + pt, nnz - 1))
    bconns = [bD, bR, bT, bL, bF, bB]
    zoff = np.repeat(np.arange(nnz) * lay, nnx)
    zoffy = np.repeat(np.arange(nnz) * lay, nny)
    surf = [np.tile(np.arange(nnx), nnz) + zoff,
            np.tile(nnx * np.arange(1, nny + 1) - 1, nnz) + zoffy,
            np.tile(nnx * (nny - 1) + np.arange(nnx), nnz) + zoff,
            np.tile(nnx * np.arange(nny), nnz) + zoffy,
            np.arange(lay),
            lay * (nnz - 1) + np.arange(lay)]
    surf = [x.astype(np.int64) for x in surf]
    normals = np.array([[0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, -1], [0, 0, 1]])
    nvecs = [np.tile(normals[f], (bconns[f].shape[0], 1)) for f in range(6)]
    return coords, nconn, surf, bconns, nvecs


def mesh_func_for(ndim, periodic, shape_func_ord=1):
    if ndim == 2 and shape_func_ord == 2:
        return create_periodic_mesh_2d_quad9 if periodic else create_mesh_2d_quad9
    if ndim == 2:
        return create_periodic_mesh_2d if periodic else create_mesh_2d
    if shape_func_ord == 2:
        return create_mesh_3d_hex27
    return create_mesh_3d


def create_mesh_2d_quad9(nex, ney, lx, ly):
    """Rectangular quad9 mesh.  regular_mesh.py:107-145.

    Node grid (2 nex + 1) x (2 ney + 1) at half spacing; element (ie, je) -> je*nex + ie with
    local nodes [corner (0,0), mid (1,0), corner (2,0), mid (2,1), corner (2,2), mid (1,2),
    corner (0,2), mid (0,1), centre (1,1)] in node-grid offsets from (2 ie, 2 je).  Faces
    [down, right, top, left]; bconns are consecutive node triples along each face.
    """
    nnx, nny = 2 * nex + 1, 2 * ney + 1
    dx, dy = lx / nex, ly / ney
    i = np.arange(nnx)
    j = np.arange(nny)
    coords = np.stack(np.broadcast_arrays(i[None, :] * dx / 2, j[:, None] * dy / 2), axis=-1).reshape(-1, 2)
    e = np.arange(nex * ney)
    ie, je = e % nex, e // nex
    off = [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (0, 1), (1, 1)]
    nconn = np.stack([(2 * ie + a) + (2 * je + b) * nnx for a, b in off], axis=1).astype(np.int64)
    surf = [np.arange(nnx, dtype=np.int64),
            np.arange(nnx - 1, nnx * nny, nnx, dtype=np.int64),
            np.arange(nnx * (nny - 1), nnx * nny, dtype=np.int64),
            np.arange(0, nnx * nny, nnx, dtype=np.int64)]
    bconns = [np.stack([s_[:-2:2], s_[1:-1:2], s_[2::2]], axis=1) for s_ in surf]
    normals = np.array([[0, -1], [1, 0], [0, 1], [-1, 0]], dtype=np.int64)
    nvecs = [np.tile(normals[f], (bconns[f].shape[0], 1)) for f in range(4)]
    return coords, nconn, surf, bconns, nvecs


def create_periodic_mesh_2d_quad9(nex, ney, lx, ly):
    """Doubly periodic quad9 mesh, node grid 2 nex x 2 ney (no duplicated nodes).
    regular_mesh.py:182-218."""
    nnx, nny = 2 * nex, 2 * ney
    dx, dy = lx / nex, ly / ney
    i = np.arange(nnx)
    j = np.arange(nny)
    coords = np.stack(np.broadcast_arrays(i[None, :] * dx / 2, j[:, None] * dy / 2), axis=-1).reshape(-1, 2)
    e = np.arange(nex * ney)
    ie, je = e % nex, e // nex
    off = [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (0, 1), (1, 1)]
    nconn = np.stack([(2 * ie + a) % nnx + ((2 * je + b) % nny) * nnx for a, b in off], axis=1).astype(np.int64)
    return coords, nconn, None, None, None


def simple_mesh(num_elem, domain_size, periodic, dirichlet_faces, outlet_faces,
                partial_dirichlet=None, shape_func_ord=1):
    """Mesh + BC index sets.  simulation.py:96-183.

    gamma_d: sorted unique Dirichlet nodes tiled per component (+nn*i);
    U_d zeros; bconn/nvec concatenated over Dirichlet faces; gamma_p: unique
    outlet nodes, else the scalar 0.
    """
    ndim = len(num_elem)
    mesh_func = mesh_func_for(ndim, periodic, shape_func_ord)
    coords, nconn, surf, bconns, nvecs = mesh_func(*num_elem, *domain_size)
    nn = coords.shape[0]
    h = domain_size[0] / num_elem[0]
    bconn = nvec = gamma_d = U_d = None
    dirichlet_nodes, outlet_nodes, bconn_list, nvec_list = [], [], [], []
    if surf is not None:
        surf = list(surf)
    for f in dirichlet_faces or []:
        dirichlet_nodes.append(surf[f])
        bconn_list.append(bconns[f])
        nvec_list.append(nvecs[f])
    for f in outlet_faces or []:
        outlet_nodes.append(surf[f])
    if partial_dirichlet:
        for f, (s_frac, e_frac, rem) in partial_dirichlet.items():
            nodes = surf[f]
            nfe = bconns[f].shape[0]
            s_idx = round(s_frac * nfe)
            e_idx = round(e_frac * nfe)
            bconn_list.append(bconns[f][s_idx:e_idx])
            nvec_list.append(nvecs[f][s_idx:e_idx])
            if shape_func_ord == 2:
                s_idx, e_idx = 2 * s_idx, 2 * e_idx
            surf[f] = nodes[s_idx:e_idx + 1]
            dirichlet_nodes.append(nodes[s_idx:e_idx + 1])
            if rem == "outlet":
                outlet_nodes.append(nodes[:s_idx])
                outlet_nodes.append(nodes[e_idx + 1:])
    if dirichlet_nodes:
        g = np.unique(np.concatenate(dirichlet_nodes))
        gamma_d = np.concatenate([g + nn * i for i in range(ndim)]).astype(np.int64)
        U_d = np.zeros(gamma_d.shape[0])
        bconn = np.concatenate(bconn_list, axis=0)
        nvec = np.concatenate(nvec_list, axis=0)
    if outlet_nodes:
        gamma_p = np.unique(np.concatenate(outlet_nodes)).astype(np.int64)
    else:
        gamma_p = 0
    return coords, nconn, surf, nn, h, bconn, nvec, gamma_d, U_d, gamma_p, mesh_func


def set_face_velocity(U_d, nodes, gamma_d, values, nn, var):
    """simulation.py:219-226 (searchsorted into the sorted gamma_d)."""
    idx = np.searchsorted(gamma_d, np.asarray(nodes) + nn * (var - 1))
    out = np.array(U_d, dtype=float, copy=True)
    out[idx] = values
    return out
