from . import regular_mesh
from . import simulation
