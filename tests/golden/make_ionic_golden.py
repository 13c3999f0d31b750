"""Write single-cell golden traces of the reference's TTP and CRN ionic models.

These two models (tables + differentiate) are not fixed by any closed form
or by a golden vector in the reference, so their input/output vectors are
produced once, in the development container, by the reference classes
themselves (torchcor/ionic/TenTusscherPanfilov.py, CourtemancheRamirezNattel.py,
which need only torch and numpy).  The output is data only: an ``.npz`` per
case with the protocol, V trace and final states.

Run from the repository root:  python tests/golden/make_ionic_golden.py
Requires the reference checkout at /root/reference (not present at test time).
"""
import json
import sys
from pathlib import Path

import numpy as np

REF = "/root/reference"
OUT = Path(__file__).resolve().parent

# Single-cell protocol: forward Euler on V with the model's own state update,
# a stimulus of STIM (uA/uF) for t in [0, 2] ms, t accumulated as a double.
DT = 0.01
STIM = 40.0
KEEP = 10          # keep every 10th V sample

CASES = [
    ("ttp_endo", "TenTusscherPanfilov", dict(cell_type="ENDO"), 40000),
    ("ttp_epi", "TenTusscherPanfilov", dict(cell_type="EPI"), 40000),
    ("crn", "CourtemancheRamirezNattel", dict(), 40000),
]
STATES = {
    "TenTusscherPanfilov": ["CaSR", "CaSS", "Cai", "D", "F", "F2", "FCaSS", "H", "J", "Ki",
                            "M", "Nai", "R", "R_", "S", "Xr1", "Xr2", "Xs"],
    "CourtemancheRamirezNattel": ["Ca_rel", "Ca_up", "Cai", "Ki", "d", "f", "f_Ca", "h", "j", "m",
                                  "oa", "oi", "u", "ua", "ui", "v", "w", "xr", "xs"],
}


def run(cls, kwargs, nsteps, dtype):
    import torch
    model = cls(dt=DT, device=torch.device("cpu"), dtype=dtype, **kwargs)
    V = model.initialize(1)
    trace = [V.item()]
    t = 0.0
    for step in range(1, nsteps + 1):
        dV = model.differentiate(V)
        V = V + DT * dV
        t += DT
        if 0.0 <= t <= 2.0:
            V = V + DT * STIM
        if step % KEEP == 0:
            trace.append(V.item())
    states = {k: getattr(model, k).item() for k in STATES[cls.__name__]}
    return np.array(trace), states


def main():
    sys.path.insert(0, REF)
    import torch
    import torchcor.ionic as ref_ionic
    meta = {"dt": DT, "stim": STIM, "stim_window": [0.0, 2.0], "keep_every": KEEP,
            "torch": torch.__version__, "cpu_capability": torch.backends.cpu.get_cpu_capability()}
    for name, clsname, kwargs, nsteps in CASES:
        cls = getattr(ref_ionic, clsname)
        for dtype, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
            trace, states = run(cls, kwargs, nsteps, dtype)
            np.savez_compressed(OUT / f"ionic_{name}_{tag}.npz", V=trace,
                                state_names=np.array(list(states)),
                                state_values=np.array(list(states.values())),
                                nsteps=nsteps)
            print(name, tag, trace.min(), trace.max(), flush=True)
    (OUT / "ionic_golden_meta.json").write_text(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
