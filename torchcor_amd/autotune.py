"""Autotuner for the block shapes of the SELL/PCG passes.

Three shapes decide how rows map onto blocks in the bandwidth-bound passes:

* ``pcg_slices_per_block``: the single-run PCG passes (and the time-step
  runtime) run on G = min(ceil(slices / s), 1024) blocks of eight
  wavefronts, a slice being 64 rows.  Default s = 8 (one slice per wavefront
  per block round).
* ``serve_slices_per_block``: a system in the served pool gets
  min(ceil(slices / s), 1024) blocks of its own.  Default s = 16.  Tuned as
  one value for all systems, on a replay of a seeded mixed-size stream.
* ``batch_max_blocks``: the cap on the blocks of the shared-mesh batch's
  passes.  Default 1024.

Each shape groups the PCG partial sums, so the last bits of a run depend on
it; every shape is chosen from the size of the run's own system alone, so a
served run still gives the same bits alone or in any mix.

``python -m torchcor_amd.autotune --out FILE`` measures every candidate on
synthetic cubes (the benchmark's TTP workload) of each size class on the
current GPU, each candidate on a fresh run over the same step window (the
PCG iteration count changes as the wave spreads), in interleaved rounds,
and writes the fastest per class (median over the rounds).  The engines
read TABLE_PATH (committed, measured on an MI355X); TORCHCOR_AMD_TUNING=0
or ``set_enabled(False)`` turns the table off (defaults everywhere).
"""
import argparse
import json
import os
import time
from pathlib import Path

TABLE_PATH = Path(__file__).resolve().parent / "tuning_mi355x.json"

CANDIDATES = {
    "pcg_slices_per_block": (8, 16, 32),
    "serve_slices_per_block": (8, 16, 32),
    "batch_max_blocks": (256, 512, 1024),
}
DEFAULTS = {"pcg_slices_per_block": 8, "serve_slices_per_block": 16, "batch_max_blocks": 1024}

_enabled = os.environ.get("TORCHCOR_AMD_TUNING", "1") != "0"
_table = None


def set_enabled(on):
    """Use the tuning table (True) or the default shapes everywhere (False)."""
    global _enabled
    _enabled = bool(on)


def load_table(path=None):
    """The tuning table: {knob: [{"max_rows": R, "value": v}, ...]} sorted by R."""
    p = Path(path) if path is not None else TABLE_PATH
    if not p.exists():
        return {}
    with open(p) as f:
        t = json.load(f)
    return {k: sorted(v, key=lambda e: e["max_rows"]) for k, v in t.get("knobs", {}).items()}


def lookup(knob, rows):
    """The tuned value of ``knob`` for a system of ``rows`` rows (the batch:
    rows x lanes), or its default when tuning is off or the table has no
    class for it."""
    global _table
    if not _enabled:
        return DEFAULTS[knob]
    if _table is None:
        _table = load_table()
    for e in _table.get(knob, ()):
        if rows <= e["max_rows"]:
            return int(e["value"])
    return DEFAULTS[knob]


def pcg_blocks(n, slices_per_block):
    """Blocks of the PCG passes for n rows at ``slices_per_block``: at most
    1024 and at most the default grid (ceil(slices/8), 1024 at most, a
    multiple of 8 when at least 8)."""
    ns = -(-int(n) // 64)
    default = min(-(-ns // 8), 1024)
    if default >= 8:
        default -= default % 8
    g = min(-(-ns // int(slices_per_block)), 1024, default)
    if g >= 8:
        g -= g % 8
    return max(g, 1)


# ------------------------------------------------------------------ tuner
def _median(v):
    v = sorted(v)
    return v[len(v) // 2]


def _tune_pcg(dev, sizes, rounds, steps, log):
    """Each candidate on a fresh run of the same step window (the iteration
    count grows as the wave spreads, so candidates must see the same steps)."""
    import torch
    from . import _lib as L
    from . import meshes
    from .ionic import TenTusscherPanfilov
    from .simulator import Monodomain
    out = []
    for N in sizes:
        nodes, elems, regions, fib = meshes.cube_workload(N)
        n = N ** 3
        cands = sorted({pcg_blocks(n, s): s for s in CANDIDATES["pcg_slices_per_block"]}.items())
        times = {s: [] for _, s in cands}
        for _ in range(rounds):
            for G, s in cands:
                sim = Monodomain(TenTusscherPanfilov("ENDO", 0.01, device=dev,
                                                     dtype=torch.float32),
                                 T=1000.0, dt=0.01, device=dev, dtype=torch.float32)
                sim.load_arrays(nodes, elems, regions, fib)
                sim.add_conductivity([0], 0.5272, 0.2076)
                sim.add_conductivity([1], 0.9074, 0.3332)
                sim.add_stimulus(meshes.corner_nodes(N, max(2, N // 10)), 0.0, 1.0, 100)
                sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
                L.call("tc_pcg_set_blocks", sim.system._pcg, G)
                sim.advance(200)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                sim.advance(steps)
                torch.cuda.synchronize()
                times[s].append((time.perf_counter() - t0) / steps * 1e3)
                sim.finish()
                del sim
                torch.cuda.empty_cache()
        best = min(times, key=lambda s: _median(times[s]))
        row = {"rows": n, "value": best,
               "ms_per_step": {str(s): round(_median(t), 5) for s, t in times.items()},
               "blocks": {str(s): G for G, s in cands}}
        log(f"[autotune] pcg {N}^3: {row}")
        out.append(row)
    return out


def stream_requests(n_req=24, seed=7, t_range=(20.0, 60.0)):
    """A seeded stream like the served benchmark's: cubes of 11..81 nodes a
    side (log-uniform), MS2v/TTP/CRN, T uniform in t_range, scaled
    conductivities, one boundary patch stimulus each; all present at once."""
    import numpy as np
    from . import meshes
    from .serve import Request
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n_req):
        N = int(round(float(np.exp(rng.uniform(np.log(11), np.log(81))))))
        model = ("ms2v", "ttp", "crn")[int(rng.integers(3))]
        T = float(rng.uniform(*t_range))
        sc = float(rng.uniform(0.5, 1.5))
        nodes, elems, regions, fib = meshes.cube_workload(N)
        surf = meshes.cube_surface_nodes(N)
        patch = meshes.surface_patch(nodes, surf, min(200, len(surf) // 4), rng)
        cond = [([0], 0.5272 * sc, 0.2076 * sc), ([1], 0.9074 * sc, 0.3332 * sc)]
        reqs.append(Request(nodes, elems, regions, fib, cond, model,
                            [(patch, 0.0, 1.0, 100.0)], T, tag=k))
    return reqs


def _tune_serve(dev, rounds, log, n_req=24):
    """One slices-per-block for every system, chosen on a replay of a seeded
    stream of mixed sizes (all runs admitted up front, to the end): in the
    pool the large systems keep the GPU busy, and a shape that helps a
    small system fill an idle GPU can slow the pool as a whole."""
    import torch
    from .serve import ServeEngine
    reqs = stream_requests(n_req)
    rows = sum(-(-r.nodes.shape[0] // 64) * 64 for r in reqs)
    times = {s: [] for s in CANDIDATES["serve_slices_per_block"]}
    for _ in range(rounds):
        for s in times:
            eng = ServeEngine(cap_nodes=rows, cap_entries=16 * rows, max_sys=n_req + 4,
                              dtype=torch.float32, device=dev, slices_per_block=s)
            prs = [eng.prepare(r) for r in reqs]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            eng.run_until_done(prs)
            torch.cuda.synchronize()
            times[s].append(time.perf_counter() - t0)
            eng.close()
            del eng, prs
            torch.cuda.empty_cache()
    best = min(times, key=lambda s: _median(times[s]))
    row = {"rows": 2 ** 62, "stream_requests": n_req, "value": best,
           "wall_s": {str(s): round(_median(t), 4) for s, t in times.items()}}
    log(f"[autotune] serve stream of {n_req} runs: {row}")
    return [row]


def _tune_batch(dev, sizes, rounds, steps, log):
    import numpy as np
    import torch
    from . import meshes
    from .batch import SharedMeshBatch
    from .ionic import ModifiedMS2v, TenTusscherPanfilov
    out = []
    for N in sizes:
        nodes, elems, regions, fib = meshes.cube_workload(N)
        rng = np.random.default_rng(2024)
        surf = meshes.cube_surface_nodes(N)
        specs = [(k < 32, rng.uniform(0.5, 1.5), rng.uniform(-30, 30),
                  meshes.surface_patch(nodes, surf, min(200, len(surf) // 4), rng))
                 for k in range(64)]
        times = {b: [] for b in CANDIDATES["batch_max_blocks"]}
        for _ in range(rounds):
            for mb in times:
                bat = SharedMeshBatch(T=1000.0, dt=0.01, device=dev, dtype=torch.float32)
                bat.load_arrays(nodes, elems, regions, fib)
                bat.add_conductivity([0], 0.5272, 0.2076)
                bat.add_conductivity([1], 0.9074, 0.3332)
                ms = ModifiedMS2v(0.01, device=dev, dtype=torch.float32)
                ttp = TenTusscherPanfilov("ENDO", 0.01, device=dev, dtype=torch.float32)
                for is_ms, sc, rot, patch in specs:
                    bat.add_instance(ms if is_ms else ttp, sc, rot, [(patch, 0.0, 1.0, 100.0)])
                bat.start(1e-5, 1e-5, 100, max_blocks=mb)
                bat.advance(100)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                bat.advance(steps)
                torch.cuda.synchronize()
                times[mb].append((time.perf_counter() - t0) / steps * 1e3)
                bat.finish()
                del bat
                torch.cuda.empty_cache()
        best = min(times, key=lambda b: _median(times[b]))
        row = {"rows": N ** 3 * 64, "value": best,
               "ms_per_step": {str(b): round(_median(t), 5) for b, t in times.items()}}
        log(f"[autotune] batch 64 x {N}^3: {row}")
        out.append(row)
    return out


def _classes(rows_and_values):
    """Size classes from measured (rows, value) points: each point's class
    reaches halfway (geometrically) to the next; the last is open-ended."""
    pts = sorted(rows_and_values, key=lambda e: e["rows"])
    out = []
    for k, e in enumerate(pts):
        hi = int((e["rows"] * pts[k + 1]["rows"]) ** 0.5) if k + 1 < len(pts) else 2 ** 62
        out.append({"max_rows": hi, "value": e["value"]})
    return out


def tune(device=None, quick=False, log=print):
    """Measure every knob on this GPU; returns the table (see load_table)."""
    import torch
    dev = torch.device(device) if device is not None else torch.device("cuda:0")
    rounds = 2 if quick else 3
    t0 = time.perf_counter()
    pcg = _tune_pcg(dev, (21, 41, 61, 81, 101) if not quick else (21, 61),
                    rounds, 100 if quick else 200, log)
    serve = _tune_serve(dev, rounds, log, n_req=12 if quick else 24)
    batch = _tune_batch(dev, (21, 41) if not quick else (21,), rounds, 50 if quick else 150, log)
    return {"device": torch.cuda.get_device_name(dev),
            "seconds": round(time.perf_counter() - t0, 1),
            "knobs": {"pcg_slices_per_block": _classes(pcg),
                      "serve_slices_per_block": _classes(serve),
                      "batch_max_blocks": _classes(batch)},
            "measurements": {"pcg": pcg, "serve": serve, "batch": batch}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="where to write the table (JSON)")
    ap.add_argument("--quick", action="store_true", help="fewer sizes and steps")
    args = ap.parse_args()
    set_enabled(False)            # measure from the defaults, not from an old table
    table = tune(quick=args.quick, log=lambda m: print(m, flush=True))
    table["command"] = "python -m torchcor_amd.autotune --out FILE" + (" --quick" if args.quick
                                                                        else "")
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps(table["knobs"]), flush=True)


if __name__ == "__main__":
    main()
