"""torchcor_amd: MI355X-native monodomain simulator behind torchcor's Python API.

Mirrors ``torchcor/__init__.py`` and ``torchcor/device.py`` of the reference:
dtype aliases and device selection.  The compute path lives in the HIP
library ``libtorchcor_amd.so`` (see ``_lib``); subpackages ``core``,
``ionic`` and ``simulator`` load it on import.
"""
import torch

float32 = torch.float32
float64 = torch.float64


def set_device(name=None, verbose=True):
    """Select the GPU (torchcor/device.py:3-22)."""
    if name is None:
        device = torch.device("cuda" if torch.cuda.is_available() else "cpu")
    else:
        device = torch.device(name)
    if torch.cuda.is_available():
        torch.cuda.set_device(device)
        if verbose:
            idx = torch.cuda.current_device()
            props = torch.cuda.get_device_properties(idx)
            print(f"GPU: {torch.cuda.get_device_name(idx)}", flush=True)
            print(f"Total Memory: {props.total_memory / (1024 ** 3):.2f} GB", flush=True)
    elif verbose:
        print("No GPU available. Using CPU instead.", flush=True)


def get_device():
    """Current device (torchcor/device.py:26-33)."""
    if torch.cuda.is_available():
        return torch.device(f"cuda:{torch.cuda.current_device()}")
    return torch.device("cpu")
