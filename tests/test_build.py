"""The C-ABI library loads and exports every function the header declares;
host-side logic (mesh files, stimulus schedule) works without a GPU."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_header_symbols():
    import torchcor_amd._lib as L
    header = (ROOT / "include" / "torchcor_amd.h").read_text()
    names = set(re.findall(r"^\s*(?:int|const char\*)\s+(tc_\w+)\s*\(", header, re.M))
    assert len(names) >= 20
    for nm in sorted(names):
        assert hasattr(L.lib, nm), nm
    assert set(L.EXPORTS) == names
    assert L.lib.tc_abi_version() == 3


def test_error_path_reports_text():
    import ctypes as C
    import torchcor_amd._lib as L
    nb = C.c_size_t(0)
    rc = L.lib.tc_pattern_workspace_size(0, 4, C.byref(nb))
    assert rc == -1 and b"bad arguments" in L.lib.tc_last_error()
    with pytest.raises(L.TorchcorError):
        L.call("tc_pattern_workspace_size", 10, 5, C.byref(nb))


def test_no_cpu_fallback():
    import torch
    import torchcor_amd._lib as L
    with pytest.raises(L.TorchcorError):
        L.ptr(torch.zeros(3))


def test_alias_package_surface():
    import torchcor as tc
    from torchcor.simulator import Monodomain  # noqa: F401
    from torchcor.simulator.monodomain import Monodomain as M2  # noqa: F401
    from torchcor.ionic import CourtemancheRamirezNattel, ModifiedMS2v, TenTusscherPanfilov  # noqa: F401
    assert tc.float32 is not None and callable(tc.get_device)
    m = TenTusscherPanfilov("ENDO", 0.01)
    assert m.GKs_init == 0.392 and m.Gto_init == 0.073
    m = TenTusscherPanfilov("EPI", 0.01)
    assert m.Gto_init == 0.294
    ms = ModifiedMS2v(0.01)
    ms.u_gate = 0.1
    assert ms.u_gate == 0.1 and ms.DV == 100.0


def test_alias_module_paths():
    """Every module path a reference user script imports from resolves
    (torchcor/__init__.py, torchcor/core/__init__.py:1-9,
    torchcor/ionic/__init__.py) and names the same objects."""
    import importlib
    import torchcor_amd
    from torchcor_amd import core, ionic
    paths = {
        "torchcor.device": ["get_device", "set_device"],
        "torchcor.core.assemble": ["Matrices3D", "Matrices3DSurface"],
        "torchcor.core.solver": ["ConjugateGradient"],
        "torchcor.core.preconditioner": ["Preconditioner"],
        "torchcor.core.mesh": ["MeshReader"],
        "torchcor.core.stimulation": ["Stimuli", "load_stimulus_region"],
        "torchcor.core.conductivity": ["Conductivity"],
        "torchcor.core.reorder": ["RCM"],
        "torchcor.core.visualize": ["VTK3D", "VTK3DSurface"],
        "torchcor.ionic.TenTusscherPanfilov": ["TenTusscherPanfilov"],
        "torchcor.ionic.ModifiedMS2v": ["ModifiedMS2v"],
        "torchcor.ionic.CourtemancheRamirezNattel": ["CourtemancheRamirezNattel"],
    }
    for path, names in paths.items():
        mod = importlib.import_module(path)
        for name in names:
            ours = getattr(torchcor_amd, name, None) or getattr(core, name, None) \
                or getattr(ionic, name, None)
            if ours is None:
                from torchcor_amd.core import vtk
                ours = getattr(vtk, name)
            assert getattr(mod, name) is ours, (path, name)


def test_mesh_files_round_trip(tmp_path):
    from torchcor_amd import meshes
    from torchcor_amd.core.mesh import MeshReader
    nodes, elems = meshes.kuhn_cube(4)
    regions = np.arange(len(elems)) % 3
    fib = meshes.rotating_fibres(nodes, elems)
    meshes.write_mesh(tmp_path, nodes, elems, regions, fib)
    n2, e2, r2, f2 = MeshReader(tmp_path).read(unit_conversion=1000)
    assert np.allclose(n2, nodes) and np.array_equal(e2, elems)
    assert np.array_equal(r2, regions) and np.allclose(f2, fib)


def test_mesh_count_mismatch_raises(tmp_path):
    from torchcor_amd import meshes
    from torchcor_amd.core.mesh import MeshReader
    nodes, elems = meshes.kuhn_cube(3)
    meshes.write_mesh(tmp_path, nodes, elems, np.zeros(len(elems), int), np.zeros((len(elems), 3)))
    pts = tmp_path / "mesh.pts"
    lines = pts.read_text().splitlines()
    pts.write_text("\n".join([str(len(nodes) + 1)] + lines[1:]) + "\n")
    with pytest.raises(ValueError):
        MeshReader(tmp_path).read()


def test_stimulus_schedule_and_vtx(tmp_path):
    import torch
    from torchcor_amd import meshes
    from torchcor_amd.core.stimulation import Stimuli, load_stimulus_region
    p = meshes.write_vtx(tmp_path / "s.vtx", [1, 3, 5])
    assert load_stimulus_region(p).tolist() == [1, 3, 5]
    bad = tmp_path / "bad.vtx"
    bad.write_text("4\nextra\n1\n2\n")
    with pytest.raises(Exception):
        load_stimulus_region(bad)
    s = Stimuli(10, torch.device("cpu"), torch.float32)
    s.add(p, 0.0, 1.0, 100.0, period=50.0)
    s.add([0, 1], 5.0, 1.0, 30.0, period=50.0)
    with pytest.raises(Exception):
        s.add(p, 0.0, 1.0, 1.0, period=None)
    assert s.mask.tolist() == [2, 3, 0, 1, 0, 1, 0, 0, 0, 0]
    t, on = 0, []
    for n in range(1, 700):
        t += 0.01
        on.append(s.active(t))
    assert on[0] == 1 and on[98] == 1 and on[99] == 0 and on[500] == 2 and on[599] == 2
    assert s.apply(0.5).tolist() == [0, 100, 0, 100, 0, 100, 0, 0, 0, 0]


def test_ctypes_structs_match_c_layout(tmp_path):
    """Every ctypes mirror of a header struct has the C size and field offsets."""
    import ctypes as C
    import subprocess
    import torchcor_amd._lib as L
    pairs = {"tc_sell": L.Sell, "tc_ionic_io": L.IonicIO, "tc_ms2v_params": L.MS2vParams,
             "tc_ttp_params": L.TTPParams, "tc_crn_params": L.CRNParams,
             "tc_stimulus": L.Stimulus, "tc_sim_desc": L.SimDesc,
             "tc_batch_group": L.BatchGroup, "tc_batch_desc": L.BatchDesc,
             "tc_serve_desc": L.ServeDesc, "tc_serve_system": L.ServeSystem}
    lines = ['#include <stdio.h>', '#include <stddef.h>',
             f'#include "{ROOT / "include" / "torchcor_amd.h"}"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                         text=True).stdout.splitlines())
    for cname, py in pairs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_sigma_split_reassembles_rotated_tensor():
    """The scalar/vector/tensor split of the conductivity tensor reproduces
    sigma_t I + (sigma_l - sigma_t) f' f'^T for fibres rotated about z."""
    import math
    import torch
    from torchcor_amd.batch import rotate_z, sigma_components
    rng = np.random.default_rng(5)
    E = 50
    f = rng.normal(size=(E, 3))
    f /= np.linalg.norm(f, axis=1, keepdims=True)
    sl, st = rng.uniform(0.2, 1.0, E), rng.uniform(0.05, 0.2, E)
    S = [c.numpy() for c in sigma_components(torch.tensor(sl), torch.tensor(st), torch.tensor(f))]
    for deg in (0.0, 17.0, -30.0, 90.0):
        fr = rotate_z(f, deg)
        ref = st[:, None, None] * np.eye(3) + (sl - st)[:, None, None] * np.einsum("ei,ej->eij", fr, fr)
        d = math.radians(deg)
        got = S[0] + math.cos(d) * S[1] + math.sin(d) * S[2] + math.cos(2 * d) * S[3] + math.sin(2 * d) * S[4]
        assert np.abs(got - ref).max() < 1e-14


def test_bench_trace_summary(tmp_path):
    """bench.py --trace-summary: a pass is a run of time-step kernels; its
    window starts at the K-th last ionic launch that ran (short launches
    of gated steps do not count) and takes every kernel after it."""
    import subprocess
    import sys
    hdr = '"Kind","Kernel_Name","Start_Timestamp","End_Timestamp"\n'
    t = [1000]
    lines = []

    def k(name, us):
        lines.append(f'"KERNEL_DISPATCH","{name}",{t[0]},{t[0] + us * 1000}\n')
        t[0] += us * 1000 + 500

    ttp = "void tc::(anonymous namespace)::k_ttp<float, true, true>(long, float*)"
    step = [("void tc::(anonymous namespace)::k_rhs_init<float, double>(long)", 30),
            ("void tc::(anonymous namespace)::k_pcg_spmv<float, short>(long)", 20),
            ("void tc::(anonymous namespace)::k_pcg_update<float, double, true>(long)", 12)]
    k("void tc::k_local_tet(int)", 100)                    # assembly: not a step kernel
    for s in range(6):                                      # 2 warm-up + 4 timed
        k(ttp, 40)
        for nm, us in step:
            k(nm, us)
    k(ttp, 10)                                              # a gated launch
    k("void tc::(anonymous namespace)::k_pcg_update<float, double>(long)", 9)
    k("__amd_rocclr_copyBuffer", 3)                         # ends the pass
    f = tmp_path / "kt_kernel_trace.csv"
    f.write_text(hdr + "".join(lines))
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--trace-summary", str(f),
                          "--steps", "4"], capture_output=True, text=True, check=True).stdout
    rows = {ln.split()[0]: ln.split() for ln in out.splitlines() if ln.startswith("k_")}
    assert rows["k_ttp"][1] == "5" and rows["k_rhs_init"][1] == "4"
    assert rows["k_pcg_update+check"][1] == "4" and rows["k_pcg_update"][1] == "1"
    assert float(rows["k_ttp"][4]) == 10.0 and float(rows["k_rhs_init"][2]) == 30.0
    assert out.startswith("pass 0: last 4 steps")
