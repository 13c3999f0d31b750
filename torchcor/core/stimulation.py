"""torchcor.core.stimulation (torchcor/core/stimulation.py) resolved to the MI355X package."""
from torchcor_amd.core.stimulation import *  # noqa: F401,F403
