"""torchcor.core.assemble (torchcor/core/assemble.py) resolved to the MI355X package."""
from torchcor_amd.core.assemble import *  # noqa: F401,F403
