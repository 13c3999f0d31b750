from torchcor_amd.core import *  # noqa: F401,F403
from torchcor_amd.core import __all__  # noqa: F401
