"""Simulation drivers (mirrors torchcor/simulator of the reference)."""
from .monodomain import Monodomain

__all__ = ["Monodomain"]
