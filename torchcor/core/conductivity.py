"""torchcor.core.conductivity (torchcor/core/conductivity.py) resolved to the MI355X package."""
from torchcor_amd.core.conductivity import *  # noqa: F401,F403
