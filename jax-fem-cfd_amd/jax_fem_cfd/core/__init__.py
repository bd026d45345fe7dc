from . import element_calc
from . import element_mult
from . import multigrid
from . import shape_functions
