"""torchcor.core.reorder (torchcor/core/reorder.py) resolved to the MI355X package."""
from torchcor_amd.core.reorder import *  # noqa: F401,F403
