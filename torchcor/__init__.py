"""Import-compatible alias of ``torchcor_amd`` under the reference's package name.

``import torchcor as tc``, ``from torchcor.simulator import Monodomain`` and
``from torchcor.ionic import TenTusscherPanfilov`` resolve to the MI355X
implementation, so user scripts written for the reference run unchanged.
"""
from torchcor_amd import float32, float64, get_device, set_device  # noqa: F401
