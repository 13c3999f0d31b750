"""Ionic cell models (mirrors torchcor/ionic of the reference)."""
from .base import IonicModel
from .models import CourtemancheRamirezNattel, ModifiedMS2v, TenTusscherPanfilov

__all__ = ["IonicModel", "ModifiedMS2v", "TenTusscherPanfilov", "CourtemancheRamirezNattel"]
