"""torchcor.core.preconditioner (torchcor/core/preconditioner.py) resolved to the MI355X package."""
from torchcor_amd.core.preconditioner import *  # noqa: F401,F403
