"""Common machinery of the ionic model classes.

A model keeps its parameters as plain float attributes (settable before
``initialize``, as in the reference's demos), its state as one (S, n) tensor
on the device, and its lookup tables as device tensors.  ``differentiate``
and the fused monodomain step both run the model's HIP kernel.
"""
import ctypes as C

import torch

from .. import _lib as L


def default_device():
    if torch.cuda.is_available():
        return torch.device(f"cuda:{torch.cuda.current_device()}")
    return torch.device("cpu")


class IonicModel:
    name = "IonicModel"
    STATES = ()        # state names in kernel order
    INIT = {}          # state name -> initial value
    PARAMS = ()        # C struct field names
    V_init = 0.0

    def __init__(self, dt, device=None, dtype=torch.float32):
        object.__setattr__(self, "_states", None)
        self.dt = dt
        self.device = torch.device(device) if device is not None else default_device()
        self.dtype = dtype

    # state rows are exposed as attributes (model.Cai -> (n,) view)
    def __getattr__(self, name):
        st = object.__getattribute__(self, "_states")
        names = type(self).STATES
        if st is not None and name in names:
            return st[names.index(name)]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        st = self.__dict__.get("_states")
        if st is not None and name in type(self).STATES:
            st[type(self).STATES.index(name)].copy_(torch.as_tensor(value, dtype=st.dtype))
            return
        object.__setattr__(self, name, value)

    @property
    def states(self):
        return self._states

    def _param_values(self):
        return {f: float(getattr(self, f)) for f in self.PARAMS}

    def _alloc_states(self, n, stride=None):
        """(S, n) states; with ``stride`` the rows lie ``stride`` entries apart
        in one (S, stride) block and ``states`` is the (S, n) view."""
        if self.device.type != "cuda":
            raise L.TorchcorError(f"{self.name} runs on the GPU only (no CPU path)")
        stride = n if stride is None else int(stride)
        buf = torch.empty((len(self.STATES), stride), dtype=self.dtype, device=self.device)
        st = buf[:, :n]
        for k, s in enumerate(self.STATES):
            st[k].fill_(self.INIT[s])
        object.__setattr__(self, "_states", st)
        self.n_nodes = n

    def initialize(self, n_nodes):
        # state rows padded to whole 64-entry groups (256 bytes in float32),
        # so every row starts on a cache-line boundary
        self._alloc_states(n_nodes, stride=-(-n_nodes // 64) * 64)
        self.construct_tables()
        return torch.full((n_nodes,), self.V_init, dtype=self.dtype, device=self.device)

    def construct_tables(self):
        pass

    def _io(self, V, w=None, dV=None, Cm=0.0, dt_rhs=None, stim_mask=None, stim_bits=0,
            stim_amp=(), AT=None, RT=None, lat_t=0.0):
        io = L.IonicIO()
        io.n = V.shape[0]
        io.dt = float(self.dt)
        io.dt_rhs = float(self.dt if dt_rhs is None else dt_rhs)
        io.Cm = float(Cm)
        io.V = L.ptr(V)
        io.w = L.ptr(w)
        io.dV = L.ptr(dV)
        io.stim_mask = L.ptr(stim_mask) if stim_bits else None
        io.stim_active = int(stim_bits) & 0xFFFFFFFF
        for k, a in enumerate(stim_amp):
            io.stim_amp[k] = float(a)
        io.AT = L.ptr(AT)
        io.RT = L.ptr(RT)
        io.lat_t = float(lat_t)
        io.state_stride = self._states.stride(0) if self._states is not None else 0
        return io

    def step(self, io):
        """Run the model kernel once with a prepared ``IonicIO``."""
        raise NotImplementedError

    def differentiate(self, V):
        """dV/dt = -Iion; advances the states by dt (reference protocol)."""
        if self._states is None:
            raise RuntimeError("initialize() first")
        V = V.to(self.dtype).contiguous()
        dV = torch.empty_like(V)
        self.step(self._io(V, dV=dV))
        return dV
