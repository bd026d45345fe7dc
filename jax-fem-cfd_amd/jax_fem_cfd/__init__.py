"""MI355X (gfx950) drop-in for the matrix-free Navier-Stokes timestep of
jax-fem-cfd.

Same package layout and call surface as the reference (``jax_fem_cfd.setup``,
``.solvers``, ``.core``, ``.utils``); the compute runs in hand-written HIP
kernels of ``libfcfd.so`` through the C ABI of ``include/fcfd.h``.  Arrays
returned by the solver are float64 torch CUDA tensors (device resident).
"""
from . import core
from . import setup
from . import solvers
from . import utils
