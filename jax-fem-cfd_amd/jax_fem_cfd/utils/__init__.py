from . import jax_iterative_solvers
from . import running
