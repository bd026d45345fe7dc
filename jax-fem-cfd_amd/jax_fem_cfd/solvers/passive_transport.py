"""Navier-Stokes step plus passive scalar transport
(jax_fem_cfd/solvers/passive_transport.py).  Not built in this version: the
scalar BiCGSTAB and SUPG kernels come after the standard timestep."""


def setup_solver(node_coords, nconn, bconn, nvec, gamma_d, gamma_p, nn, gamma_d_phi, mesh_func, config):
    raise NotImplementedError("passive scalar transport is not built yet")
