"""torchcor.device (torchcor/device.py) resolved to the MI355X package."""
from torchcor_amd import get_device, set_device  # noqa: F401
