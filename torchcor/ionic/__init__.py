from torchcor_amd.ionic import *  # noqa: F401,F403
from torchcor_amd.ionic import __all__  # noqa: F401
