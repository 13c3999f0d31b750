from torchcor_amd.simulator import Monodomain  # noqa: F401
