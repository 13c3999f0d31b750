"""Numerics of the monodomain path (mirrors torchcor/core of the reference)."""
from .assemble import Matrices3D, Matrices3DSurface
from .conductivity import Conductivity
from .mesh import MeshReader
from .preconditioner import Preconditioner
from .reorder import RCM
from .solver import ConjugateGradient
from .stimulation import Stimuli, load_stimulus_region
from .system import P1System, csr_system

__all__ = ["Matrices3D", "Matrices3DSurface", "Conductivity", "MeshReader", "Preconditioner", "RCM",
           "ConjugateGradient", "Stimuli", "load_stimulus_region", "P1System", "csr_system"]
