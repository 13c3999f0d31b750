"""Stimulus schedule (torchcor/core/stimulation.py:17-63).

Each stimulus is a node set, a window [start, start+duration] inside each
period, a repeat count and an intensity.  On the device the node sets are one
32-bit mask per node (bit k = node in stimulus k); per step the host decides
which stimuli are on (the same double-precision test as the reference) and
the ionic kernel adds their intensities.
"""
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

MAX_STIMULI = 32


@dataclass
class StimConfig:
    vtx_filepath: object
    start: float
    duration: float
    intensity: float
    period: float
    count: float


def load_stimulus_region(vtx_filepath, device=None):
    """Node ids of a ``.vtx`` file: count line, a tag line, one id per line."""
    from .mesh import read_table
    with Path(vtx_filepath).open("r") as f:
        size = int(f.readline().strip())
    region = read_table(vtx_filepath, 2, dtype=np.int64).reshape(-1)
    if len(region) != size:
        raise Exception(f"Error loading {vtx_filepath}")
    return torch.from_numpy(region).to(dtype=torch.long, device=device)


class Stimuli:
    def __init__(self, n_nodes, device, dtype):
        self.n_nodes = n_nodes
        self.device = device
        self.dtype = dtype
        self.stimulus_list = []
        self.mask = torch.zeros(n_nodes, dtype=torch.int32, device=device)
        self.Istim = torch.zeros(n_nodes, dtype=dtype, device=device)

    def load_stimulus_region(self, vtx_filepath):
        return load_stimulus_region(vtx_filepath, self.device)

    def add(self, vtx_filepath, start, duration, intensity, period=None, count=1):
        """Add a stimulus; ``vtx_filepath`` may also be an array of node ids."""
        if period is None:
            raise Exception("period is unspecified.")
        k = len(self.stimulus_list)
        if k >= MAX_STIMULI:
            raise ValueError(f"at most {MAX_STIMULI} stimuli are supported")
        if isinstance(vtx_filepath, (str, Path)):
            region = self.load_stimulus_region(vtx_filepath)
        else:
            region = torch.as_tensor(np.asarray(vtx_filepath), dtype=torch.long, device=self.device)
        bit = torch.tensor(1 << k if k < 31 else -(1 << 31), dtype=torch.int32, device=self.device)
        self.mask[region] |= bit
        self.stimulus_list.append(StimConfig(vtx_filepath, start, duration, intensity, period, count))

    def active(self, t):
        """Bit set of the stimuli on at time t (stimulation.py:51-63), as an int."""
        bits = 0
        for k, s in enumerate(self.stimulus_list):
            p_t = t % s.period
            t_t = t / s.period
            if t_t <= s.count and s.start <= p_t <= s.start + s.duration:
                bits |= 1 << k
        return bits

    def amplitudes(self):
        return [float(s.intensity) for s in self.stimulus_list]

    def apply(self, t):
        """Dense stimulus current at time t (reference API)."""
        bits = self.active(t)
        if bits == 0:
            return self.Istim
        out = torch.zeros(self.n_nodes, dtype=self.dtype, device=self.device)
        for k, s in enumerate(self.stimulus_list):
            if bits >> k & 1:
                on = ((self.mask.to(torch.int64) >> k) & 1).to(self.dtype)
                out = out + on * s.intensity
        return out
