"""torchcor.ionic.CourtemancheRamirezNattel (torchcor/ionic/CourtemancheRamirezNattel.py) resolved to the MI355X package."""
from torchcor_amd.ionic import CourtemancheRamirezNattel  # noqa: F401
