"""torchcor.core.mesh (torchcor/core/mesh.py) resolved to the MI355X package."""
from torchcor_amd.core.mesh import *  # noqa: F401,F403
