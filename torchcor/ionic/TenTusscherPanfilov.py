"""torchcor.ionic.TenTusscherPanfilov (torchcor/ionic/TenTusscherPanfilov.py) resolved to the MI355X package."""
from torchcor_amd.ionic import TenTusscherPanfilov  # noqa: F401
