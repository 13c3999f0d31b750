from torchcor_amd.simulator.monodomain import Monodomain  # noqa: F401
