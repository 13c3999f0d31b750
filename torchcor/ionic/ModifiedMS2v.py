"""torchcor.ionic.ModifiedMS2v (torchcor/ionic/ModifiedMS2v.py) resolved to the MI355X package."""
from torchcor_amd.ionic import ModifiedMS2v  # noqa: F401
