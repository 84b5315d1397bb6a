from .mpo import MPO
from .mps import MPS
