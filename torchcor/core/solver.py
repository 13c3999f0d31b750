"""torchcor.core.solver (torchcor/core/solver.py) resolved to the MI355X package."""
from torchcor_amd.core.solver import *  # noqa: F401,F403
