from . import passive_transport
from . import standard
