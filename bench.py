"""Benchmark of the monodomain time step on one MI355X.

Workload (default): the 1M-node TTP cube -- 101^3 nodes (1,030,301),
6,000,000 Kuhn tetrahedra, edge 0.1 mm, TenTusscherPanfilov ENDO, float32
storage with float64 PCG reductions, dt = 0.01 ms, fibres rotating -60..+60
degrees about z, two conductivity regions split at x = 5 mm, 1,000-node
corner stimulus (100 uA/cm^2, 0..1 ms), Jacobi-PCG atol = rtol = 1e-5,
max_iter 100, linear guess on.  Synthetic mesh and inputs; random state is
not used.

A step is one time step of the whole path (ionic kernel, RHS, PCG).  W
warm-up steps run first (they include the stimulus onset); K steps are timed
between synchronisations.  value = nodes * K / elapsed (node-timesteps/s).

Per-kernel durations come from HIP events attached to the kernel launches
on the stream they run on.  Events on a launch add a few microseconds of
gap after it, so the timed pass carries them on every 16th step only; when
that gives fewer than 16 samples (short windows), a second pass over the
same step window (same warm-up, same K steps) times every step and supplies
the kernel figures.  A steady-state window (2000 steps after 2 ms) is
reported beside the headline, and the headline window run again with each
step's fixed launches replayed as a captured HIP graph (graph_replay).  Also: the batched and served forms, the
irregular shell, and the CPU baseline (a torch-CPU restatement of the step
on the host's threads, float32 and float64).
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "node-timesteps/s + PCG HBM GB/s (1M-node TTP); sims/hour & p99 lat batched"
HBM_PEAK_GBPS = 8000.0


def alg_bytes(n, nnz, S, col_bytes, xb=4, wb=4):
    """Algorithmic HBM bytes per launch of each kernel (each array read or
    written once; slice pointers ignored): float32 values and states,
    `col_bytes` of column data per SELL pass (see column_bytes), `xb`-byte
    iterate x and `wb`-byte w (8 in float32 runs with the float64 iterate and
    right-hand side)."""
    return {
        # V (= x) and S states in; S states, w out; AT, RT read
        "ionic": n * (xb + 8 * S + wb + 8),
        # M, K values + columns; w, x, x_prev, Dinv in; x_next, x_prev_next, r out
        "rhs_init": 8 * nnz + col_bytes + n * (wb + xb + 4 + 4 + xb + 4 + 4),
        # A values + columns; r, Dinv, p in; p', Ap out
        "pcg_spmv": 4 * nnz + col_bytes + 20 * n,
        # p', Ap, Dinv, r, x in; x, r out
        "pcg_update": n * (12 + 4 + xb + xb + 4),
    }


def column_bytes(system):
    """Column data one SELL pass reads: the int16 (or int32) offsets of the
    rows in per-row slices, and one int32 pattern entry per entry position of
    each shared slice (whose rows share one stencil)."""
    import torch
    ib = 2 if system.sell.dcol is not None else 4
    deg = system.row_ptr.diff().to(torch.int64)
    if system.sell.dcol is None or system.slice_tag is None:
        return ib * int(deg.sum().item())
    ns = (system.n + 63) // 64
    sh = (system.slice_tag[:ns] & 1).to(torch.bool)
    row_sh = sh.repeat_interleave(64)[: system.n]
    width = system.slice_ptr.diff().to(torch.int64) // 64
    return ib * int(deg[~row_sh].sum().item()) + 4 * int(width[sh].sum().item())


# Memory bytes per dispatch of each kernel on the default workload (101^3 TTP
# cube, float32 storage, float64 iterate and right-hand side, shared slices
# for nearly all of the cube's rows): what the L2 caches fetch from and write to the memory fabric
# (HBM, or the last-level cache in front of it), from two rocprofv3 --pmc
# runs of this script (FETCH_SIZE, then
# WRITE_SIZE, median KiB per dispatch) scaled by what the same counters read
# for a plain 256 MiB device copy (--copy-probe-mib 256): FETCH_SIZE x2.0,
# WRITE_SIZE x1.0.  profiles/cube101_ttp_step_hbm_traffic.md holds the
# counter summary and the commands.
MEASURED_TRAFFIC_101 = {
    "ionic": 173_189_344,
    "rhs_init": 174_725_792,
    "pcg_spmv": 82_761_248,
    "pcg_update": 37_302_912,
}


def build_sim(N, dev):
    import torch
    from torchcor_amd import meshes
    from torchcor_amd.ionic import TenTusscherPanfilov
    from torchcor_amd.simulator import Monodomain
    nodes, elems, regions, fib = meshes.cube_workload(N)
    sim = Monodomain(TenTusscherPanfilov("ENDO", 0.01, device=dev, dtype=torch.float32), T=100.0,
                     dt=0.01, device=dev, dtype=torch.float32)
    sim.load_arrays(nodes, elems, regions, fib)
    sim.add_conductivity([0], 0.5272, 0.2076)
    sim.add_conductivity([1], 0.9074, 0.3332)
    sim.add_stimulus(meshes.corner_nodes(N, 10), 0.0, 1.0, 100)
    return sim, (nodes, elems, regions, fib)


def host_operators(sim, N):
    """The assembled operators of the workload in CSR order on the host (float64
    copies of the float32 values) and its stimulus, for the CPU baseline."""
    import types
    import numpy as np
    from torchcor_amd import meshes
    S = sim.system
    mask = np.zeros(S.n)
    mask[meshes.corner_nodes(N, 10)] = 100.0
    return types.SimpleNamespace(
        row_ptr=S.row_ptr.cpu().numpy(), col=S.col.cpu().numpy(), n=S.n,
        M=S.Mv.double().cpu().numpy(), K=S.Kv.double().cpu().numpy(), Cm=sim.Cm, dt=sim.dt,
        theta=sim.theta, stimuli=[(mask, 0.0, 1.0, sim.T, 1)])


def cpu_baseline(ops, N, seconds=12.0):
    """The step on the host CPU: the torch-CPU restatement (oracle/torch_cpu.py:
    torch sparse CSR products, TTP on torch tensors, float64 Jacobi-PCG, the
    reference's operation order and storage types) on torch's thread count,
    float32 then float64, each timed over as many steps (after 2 untimed) as
    fit about `seconds` seconds from t = 0.  Assembly is not timed (the
    operators are the GPU's, copied to the host)."""
    import torch
    from oracle import torch_cpu
    threads = torch.get_num_threads()
    out = {}
    for dt_ in (torch.float32, torch.float64):
        name = "f32" if dt_ == torch.float32 else "f64"
        print(f"[bench] cpu baseline ({name}, torch CPU, {threads} threads)", flush=True)
        tm = torch_cpu.Monodomain(ops, torch_cpu.TenTusscherPanfilov("ENDO", 0.01, dt_), dt_)
        u = tm.initialize()
        t = 0.0
        t0 = time.perf_counter()
        for _ in range(2):
            t += 0.01
            u, _ = tm.step(u, t)
        per = (time.perf_counter() - t0) / 2
        steps = max(3, min(400, int(seconds / max(per, 1e-6))))
        t0 = time.perf_counter()
        for _ in range(steps):
            t += 0.01
            u, _ = tm.step(u, t)
        el = time.perf_counter() - t0
        out[name] = {"value": ops.n * steps / el, "steps": steps, "seconds": round(el, 1),
                     "t_window_ms": [0.02, round(t, 2)]}
    f32, f64 = out["f32"], out["f64"]
    return {"value": f32["value"], "unit": "node-timesteps/s", "cores": threads, "kind": "port",
            "sample": f"{f32['steps']} steps (t = 0.03..{f32['t_window_ms'][1]:.2f} ms) of the "
                      f"same {N}^3 TTP workload, float32, torch CPU restatement of the step "
                      f"(oracle/torch_cpu.py) on {threads} threads, {f32['seconds']} s; "
                      f"assembly excluded",
            "cpu_model": torch_cpu.cpu_model_name(),
            "physical_cores": torch_cpu.physical_cores(),
            "float64": {"value": f64["value"], "steps": f64["steps"],
                        "seconds": f64["seconds"]}}


def batched_config(dev, steps, warmup, graph=False):
    """64 simulations sharing the 41^3 cube (68,921 nodes): 32 ModifiedMS2v
    and 32 TTP ENDO, conductivity scales U[0.5, 1.5], fibre rotations
    U[-30, 30] degrees, one seeded 200-node boundary patch stimulus each
    (100 uA/cm^2, 0..1 ms), float32, T = 100 ms.  The batch and back-to-back
    single runs of the same models are timed over the same step window;
    sims/hour extrapolates the per-step time to the 10,000 steps of a run."""
    import numpy as np
    import torch
    from torchcor_amd import meshes
    from torchcor_amd.batch import SharedMeshBatch
    from torchcor_amd.ionic import ModifiedMS2v, TenTusscherPanfilov
    from torchcor_amd.simulator import Monodomain
    N, B, T, dt = 41, 64, 100.0, 0.01
    nt = int(T / dt)
    f32 = torch.float32
    nodes, elems, regions, fib = meshes.cube_workload(N)
    rng = np.random.default_rng(2024)
    surf = meshes.cube_surface_nodes(N)
    cond = [([0], 0.5272, 0.2076), ([1], 0.9074, 0.3332)]
    bat = SharedMeshBatch(T=T, dt=dt, device=dev, dtype=f32)
    bat.load_arrays(nodes, elems, regions, fib)
    for c in cond:
        bat.add_conductivity(*c)
    ms = ModifiedMS2v(dt, device=dev, dtype=f32)
    ttp = TenTusscherPanfilov("ENDO", dt, device=dev, dtype=f32)
    patches = []
    for k in range(B):
        patch = meshes.surface_patch(nodes, surf, 200, rng)
        patches.append(patch)
        bat.add_instance(ms if k < B // 2 else ttp, rng.uniform(0.5, 1.5), rng.uniform(-30, 30),
                         [(patch, 0.0, 1.0, 100.0)])
    print(f"[bench] batch: {B} x {N}^3, warmup {warmup}", flush=True)
    bat.start(1e-5, 1e-5, 100, graph=graph)
    bat.advance(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bat.advance(steps)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    plain = {}
    for name, make in (("ms2v", lambda: ModifiedMS2v(dt, device=dev, dtype=f32)),
                       ("ttp", lambda: TenTusscherPanfilov("ENDO", dt, device=dev, dtype=f32))):
        sim = Monodomain(make(), T=T, dt=dt, device=dev, dtype=f32)
        sim.load_arrays(nodes, elems, regions, fib)
        for c in cond:
            sim.add_conductivity(*c)
        sim.add_stimulus(patches[0], 0.0, 1.0, 100.0)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9, graph=graph)
        sim.advance(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.advance(steps)
        torch.cuda.synchronize()
        plain[name] = (time.perf_counter() - t0) / steps
    t_plain = 0.5 * (plain["ms2v"] + plain["ttp"])
    sph = B * 3600.0 / (t_batch * nt)
    sph_plain = 3600.0 / (t_plain * nt)
    return {"workload": f"{B} sims on cube{N} ({N ** 3} nodes): 32 MS2v + 32 TTP ENDO, f32, "
                        f"sigma scale U[0.5,1.5], fibre rotation U[-30,30] deg, 200-node patches",
            "steps_timed": steps, "t_window_ms": [warmup * dt, (warmup + steps) * dt],
            "ms_per_step": round(t_batch * 1e3, 4),
            "node_timesteps_per_s": N ** 3 * B / t_batch,
            "sims_per_hour": round(sph, 1),
            "plain_ms_per_step": {k: round(v * 1e3, 4) for k, v in plain.items()},
            "plain_sims_per_hour": round(sph_plain, 1),
            "speedup_vs_back_to_back": round(sph / sph_plain, 2),
            "max_blocks": bat.max_blocks,
            "pcg_iterations_per_step_max": sum(bat.iterations[warmup:]) / steps + 1}


def irregular_config(dev, steps, warmup, n_target=640_000, orders=(None, "rcm", "levels", "auto")):
    """Irregular volume: the thick ellipsoidal shell (semi-axes 40/30/60 mm,
    wall 10 mm; jittered-lattice Delaunay, about 640k nodes, randomly
    permuted node numbers), four conductivity regions, transmural fibres,
    five stimulus patches (0 and 5 ms onsets), TTP ENDO float32.  Runs the
    same steps in the mesh's own (scattered) order and in locality orders:
    reverse Cuthill-McKee, Cuthill-McKee levels along a Z-order curve
    ("levels"), and the simulator's default ("auto": the level order with
    rows grouped by length inside each PCG block's window); the locality
    orders run with staged gathers (per-block column sets in LDS) when their
    sets fit."""
    import torch
    from torchcor_amd import meshes
    from torchcor_amd.ionic import TenTusscherPanfilov
    from torchcor_amd.simulator import Monodomain
    t0 = time.perf_counter()
    nodes, elems, regions, fib, cond, stims = meshes.shell_workload(n_target, seed=0)
    t_mesh = time.perf_counter() - t0
    print(f"[bench] shell: {nodes.shape[0]} nodes, {elems.shape[0]} tets in {t_mesh:.1f} s",
          flush=True)
    out = {"workload": f"ellipsoidal shell 40/30/60 mm wall 10 mm, {nodes.shape[0]} nodes, "
                       f"{elems.shape[0]} Delaunay tets, permuted node order, TTP ENDO f32, "
                       f"4 regions, 5 patches",
           "n_nodes": int(nodes.shape[0]), "n_tets": int(elems.shape[0]), "mesh_s": round(t_mesh, 1),
           "steps_timed": steps, "t_window_ms": [warmup * 0.01, (warmup + steps) * 0.01]}
    for reorder in orders:
        sim = Monodomain(TenTusscherPanfilov("ENDO", 0.01, device=dev, dtype=torch.float32),
                         T=(warmup + steps) * 0.01 + 1.0, dt=0.01, device=dev, dtype=torch.float32,
                         reorder=reorder)
        sim.load_arrays(nodes, elems, regions, fib)
        for c in cond:
            sim.add_conductivity(*c)
        for ids, start in stims:
            sim.add_stimulus(ids, start, 1.0, 100)
        t0 = time.perf_counter()
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        sim.advance(warmup)
        torch.cuda.synchronize()
        sim.set_kernel_timing(True, period=16)
        t0 = time.perf_counter()
        sim.advance(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        kt = {k: round(ms / cnt * 1e3, 2) for k, (ms, cnt) in sim.kernel_times().items() if cnt}
        key = reorder or "mesh_order"
        out[key] = {"ms_per_step": round(el / steps * 1e3, 4),
                    "node_timesteps_per_s": sim.n_nodes * steps / el,
                    "setup_s": round(t_setup, 2), "nnz": sim.system.nnz,
                    "bandwidth": sim.system.bandwidth,
                    "int16_columns": sim.system.dcol is not None,
                    "sell_padding": round(sim.system.padded / sim.system.nnz - 1, 4),
                    "pcg_blocks": int(sim.system.pcg_blocks),
                    "staged_gathers": bool(sim.system.staged),
                    "row_window": int(sim.row_window),
                    "stage_chunk": int(getattr(sim.system, "stage_chunk", 0))
                    if sim.system.staged else None,
                    "stage_max_cols": int(getattr(sim.system, "stage_max_cols", 0))
                    if sim.system.staged else None,
                    "kernel_avg_us": kt,
                    "pcg_iterations_per_step": sum(sim.iterations[warmup:]) / steps + 1}
        sim.finish()
        del sim
    for o in ("rcm", "levels", "auto"):
        if o in out and "mesh_order" in out:
            out["speedup_" + o] = round(out["mesh_order"]["ms_per_step"] / out[o]["ms_per_step"], 2)
    return out


def served_config(dev, n_req, budget_nodes=4_000_000, graph=False):
    """Served stream: Poisson arrivals (seeded, mean 20 req/s) of runs on
    cubes of 11..81 nodes a side (log-uniform), ionic model uniform over
    {MS2v, TTP ENDO, CRN}, T ~ U[50, 300] ms, conductivity scale U[0.5, 1.5],
    one seeded boundary patch stimulus (100 uA/cm^2, 0..1 ms), float32; a
    device pool of `budget_nodes` rows is the HBM budget.  The baseline runs
    the same requests back to back on the single-run path, each to the end,
    in arrival order (assembly included); its latencies follow from those
    measured run times under the same arrivals."""
    import numpy as np
    import torch
    from torchcor_amd import meshes
    from torchcor_amd.serve import MODEL_CLASSES, Request, ServeEngine
    from torchcor_amd.simulator import Monodomain
    f32 = torch.float32
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / 20.0, n_req))
    reqs = []
    for k in range(n_req):
        N = int(round(float(np.exp(rng.uniform(np.log(11), np.log(81))))))
        model = ("ms2v", "ttp", "crn")[int(rng.integers(3))]
        T = float(rng.uniform(50.0, 300.0))
        sc = float(rng.uniform(0.5, 1.5))
        nodes, elems, regions, fib = meshes.cube_workload(N)
        surf = meshes.cube_surface_nodes(N)
        patch = meshes.surface_patch(nodes, surf, min(200, len(surf) // 4), rng)
        cond = [([0], 0.5272 * sc, 0.2076 * sc), ([1], 0.9074 * sc, 0.3332 * sc)]
        reqs.append(Request(nodes, elems, regions, fib, cond, model, [(patch, 0.0, 1.0, 100.0)], T,
                            arrival=float(arrivals[k]), tag=k))
    eng = ServeEngine(cap_nodes=budget_nodes, cap_entries=16 * budget_nodes, max_sys=128, dtype=f32,
                      device=dev, graph=graph)
    # warm-up on a tiny run (kernel and table set-up outside the clock)
    eng.run_until_done([eng.prepare(Request(*meshes.cube_workload(5), reqs[0].conductivities,
                                            "ttp", [], 0.5))])
    print(f"[bench] served: {n_req} requests, {sum(r.nodes.shape[0] for r in reqs)} nodes total",
          flush=True)
    records, summary = eng.serve(reqs)
    eng.close()
    # back to back on the single-run path (estimated per request)
    est = []
    for r in reqs:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim = Monodomain(MODEL_CLASSES[r.model](0.01, dev, f32), T=r.T, dt=0.01, device=dev, dtype=f32)
        sim.load_arrays(r.nodes, r.elems, r.regions, r.fibres)
        for c in r.conductivities:
            sim.add_conductivity(*c)
        sim.add_stimulus(r.stimuli[0][0], *r.stimuli[0][1:])
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9, graph=graph)
        sim.advance(sim.nt)
        sim.finish()
        torch.cuda.synchronize()
        est.append(time.perf_counter() - t0)
        del sim
    done, lat = 0.0, []
    for r, e in zip(reqs, est):
        done = max(done, r.arrival) + e
        lat.append((done - r.arrival) / (r.T / 1000.0))
    b2b_wall = done
    b2b = {"sims_per_hour": round(n_req * 3600.0 / b2b_wall, 1),
           "latency_per_sim_s_p50": round(float(np.percentile(lat, 50)), 3),
           "latency_per_sim_s_p99": round(float(np.percentile(lat, 99)), 3),
           "wall_s": round(b2b_wall, 2)}
    return {"workload": f"{n_req} requests, Poisson 20/s (seed 5), cubes 11..81 log-uniform, "
                        f"MS2v/TTP/CRN, T U[50,300] ms, f32, pool budget {budget_nodes} rows",
            "requests": n_req, "wall_s": round(summary["wall_s"], 2),
            "assembly_s": round(summary["prepare_s"], 2),
            "engine_steps": summary["steps"], "stalls": summary["stalls"],
            "advance_calls": summary["advance_calls"],
            "advance_s": round(summary["advance_s"], 2),
            "sims_per_hour": round(summary["sims_per_hour"], 1),
            "latency_per_sim_s_p50": round(summary["latency_per_sim_s_p50"], 3),
            "latency_per_sim_s_p99": round(summary["latency_per_sim_s_p99"], 3),
            "node_timesteps": int(sum(r.nodes.shape[0] * int(r.T / 0.01) for r in reqs)),
            "back_to_back_single_runs": b2b,
            "speedup_vs_back_to_back": round(summary["sims_per_hour"] / b2b["sims_per_hour"], 2)}


def copy_ceiling(nbytes, reps=10):
    """GB/s of a plain device copy that reads nbytes/2 and writes nbytes/2
    (the library's tc_stream_copy: one 16-byte load and store per thread),
    with the 256 MB last-level cache evicted before each timed copy by a
    1 GiB read (a sum): what HBM delivers to a kernel that moves the same
    bytes once.  The eviction only reads, so the copy does not also pay for
    writing back dirty lines it displaces (an eviction by a 1 GiB copy
    leaves that much dirty data and cuts the copy's rate by about a fifth).
    Median of `reps` copies timed with events on the copy's stream."""
    import torch
    from torchcor_amd import _lib as L
    half = int(nbytes // 2) // 16 * 16
    a = torch.ones(half // 4, device="cuda:0")
    b = torch.empty_like(a)
    fa = torch.ones(1 << 28, device="cuda:0")
    s = L.stream_handle()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for r in range(reps + 2):
        fa.sum()                      # clean eviction: reads 1 GiB, writes a scalar
        e0.record()
        L.call("tc_stream_copy", a.data_ptr(), b.data_ptr(), half, s)
        e1.record()
        e1.synchronize()
        if r >= 2:
            ts.append(e0.elapsed_time(e1) / 1e3)
    ts.sort()
    t = ts[len(ts) // 2]
    del a, b, fa
    torch.cuda.empty_cache()
    return round(2 * half / t / 1e9, 1)


STEP_KERNELS = ("k_ttp", "k_crn", "k_ms2v", "k_rhs_init", "k_pcg_spmv", "k_pcg_update",
                "k_pcg_check", "k_set_int", "k_copy_gated", "k_rhs_init_staged",
                "k_pcg_spmv_staged")


def _kernel_label(name):
    """Short label of a kernel name from a rocprofv3 trace; the update pass
    that also runs the stopping test (last template argument true) is
    labelled k_pcg_update+check."""
    base = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
    head = base.split("<")[0]
    for k in sorted(STEP_KERNELS, key=len, reverse=True):
        if head.endswith(k):
            if k == "k_pcg_update" and base[base.find("<") + 1:base.rfind(">")].split(",")[-1].strip() == "true":
                return "k_pcg_update+check"
            return k
    if head.startswith("tc::"):
        head = head[4:]
    return head[-40:]


def trace_summary(path, steps):
    """Per-kernel statistics of the time-step windows in a rocprofv3
    kernel_trace.csv of this script.  A pass is a run of time-step kernels
    with no other kernel between them; its window starts at the `steps`-th
    last ionic launch that ran (launches of a stalled pipeline's gated steps
    return early and are not counted) and ends with the pass's last kernel."""
    import csv
    import statistics as st
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    names = [_kernel_label(n) for _, _, n in rows]
    step = set(STEP_KERNELS) | {"k_pcg_update+check"}
    passes, cur = [], []
    for i, n in enumerate(names):
        if n in step:
            cur.append(i)
        elif cur:
            passes.append(cur)
            cur = []
    if cur:
        passes.append(cur)
    for p, idx in enumerate(passes):
        ion = [i for i in idx if names[i] in ("k_ttp", "k_crn", "k_ms2v")]
        if len(ion) < steps:
            continue
        med = st.median(rows[i][1] - rows[i][0] for i in ion)
        ran = [i for i in ion if rows[i][1] - rows[i][0] > 0.6 * med]
        if len(ran) < steps:
            continue
        first = ran[-steps]
        win = [i for i in idx if i >= first]
        t0, t1 = rows[first][0], max(rows[i][1] for i in win)
        per, busy = {}, 0.0
        for i in win:
            d = (rows[i][1] - rows[i][0]) / 1e3
            per.setdefault(names[i], []).append(d)
            busy += d
        print(f"pass {p}: last {steps} steps, {(t1 - t0) / 1e3 / steps:.2f} us/step under the "
              f"profiler, GPU busy {min(busy * 1e3 / (t1 - t0), 1.0):.3f}")
        print(f"{'kernel':20s} {'calls':>6s} {'mean_us':>8s} {'median_us':>9s} {'min_us':>7s}")
        for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
            print(f"{k:20s} {len(v):6d} {st.mean(v):8.2f} {st.median(v):9.2f} {min(v):7.2f}")
        print()
    # the batched runs and their back-to-back single runs: every launch from
    # the first batch kernel (k_bat_*) up to the first served-engine kernel
    bat = [i for i, n in enumerate(names) if n.startswith("k_bat_")]
    if bat:
        end = next((i for i, n in enumerate(names)
                    if i > bat[0] and (n == "k_pool_ionic" or n.startswith("k_sv_"))), len(rows))
        per = {}
        for i in range(bat[0], end):
            per.setdefault(names[i], []).append((rows[i][1] - rows[i][0]) / 1e3)
        print(f"batched runs and their single-run baseline: {end - bat[0]} launches")
        print(f"{'kernel':28s} {'calls':>7s} {'mean_us':>8s} {'median_us':>9s} {'total_ms':>9s}")
        for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1]))[:16]:
            print(f"{k:28s} {len(v):7d} {st.mean(v):8.2f} {st.median(v):9.2f} {sum(v) / 1e3:9.2f}")
        print()
    # the served engine's passes (k_pool_ionic, k_sv_*): first to last of them
    sv = [i for i, n in enumerate(names) if n == "k_pool_ionic" or n.startswith("k_sv_")]
    if sv:
        t0, t1 = rows[sv[0]][0], max(rows[i][1] for i in sv)
        per, busy = {}, 0.0
        for i in range(sv[0], sv[-1] + 1):
            d = (rows[i][1] - rows[i][0]) / 1e3
            busy += d
            per.setdefault(names[i], []).append(d)
        steps = len(per.get("k_pool_ionic", []))
        print(f"served engine: {steps} engine steps, window {(t1 - t0) / 1e9:.2f} s, "
              f"GPU busy {min(busy * 1e3 / (t1 - t0), 1.0):.3f}")
        print(f"{'kernel':28s} {'calls':>7s} {'mean_us':>8s} {'total_s':>8s}")
        for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1]))[:12]:
            print(f"{k:28s} {len(v):7d} {st.mean(v):8.2f} {sum(v) / 1e6:8.3f}")


def copy_probe(mib):
    """Five copies of `mib` MiB between two device buffers: each reads and
    writes exactly mib MiB, which scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(mib * 2**18, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": a.numel() * 4}), flush=True)


def kernel_stats(kt, B):
    """{kernel: launches, average microseconds, and for the bandwidth-bound
    passes their algorithmic bytes and GB/s} from kernel_times()."""
    kernels = {}
    for k, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        avg_s = ms / cnt / 1e3
        kernels[k] = {"launches": cnt, "avg_us": round(avg_s * 1e6, 2), "total_ms": round(ms, 3)}
        if k in B:
            kernels[k].update(alg_bytes=B[k], GBps=round(B[k] / avg_s / 1e9, 1))
    return kernels


def roofline(kernels, B, N):
    """The dominant kernel (largest total time) against the HBM peak."""
    dom = max((k for k in kernels if k in B), key=lambda k: kernels[k]["total_ms"])
    a = kernels[dom]["GBps"]
    traffic = MEASURED_TRAFFIC_101.get(dom) if N == 101 else None
    return {"kernel": dom, "bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(a / HBM_PEAK_GBPS, 4), "traffic": traffic,
            "alg_bytes": B[dom]}


def pcg_rate(kernels, B):
    """The PCG phase (RHS + start pass, spmv and update passes, check) by the
    algorithmic bytes its fused passes move, over its measured time."""
    ks = ("rhs_init", "pcg_spmv", "pcg_update")
    if not all(k in kernels for k in ks):
        return None
    moved = sum(kernels[k]["launches"] * B[k] for k in ks)
    t = sum(kernels[k]["total_ms"] for k in ks + ("pcg_check",) if k in kernels) / 1e3
    return round(moved / t / 1e9, 1)


def window(N, dev, warmup, steps, period, graph=False):
    """A fresh run of the workload: `warmup` steps, then `steps` steps timed
    as a whole, with per-kernel HIP events on every `period`-th step (0:
    none).  Returns (sim, elapsed s, iterations of the timed steps, t_setup)."""
    import torch
    sim, _ = build_sim(N, dev)
    t0 = time.perf_counter()
    sim.start(1e-5, 1e-5, 100, calculate_AT_RT=True, linear_guess=True,
              snapshot_interval=1e9, verbose=False, graph=graph)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    sim.advance(warmup)
    torch.cuda.synchronize()
    if period:
        sim.set_kernel_timing(True, period=period)
    it0 = len(sim.iterations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.advance(steps)
    torch.cuda.synchronize()
    return sim, time.perf_counter() - t0, sim.iterations[it0:], t_setup


def size_sweep(sides, warmup, steps):
    """node-timesteps/s of the TTP cube workload (float32, the bench's
    stimulus and conductivities) over cube sides: the size dependence the
    reference's own volume sweep measures (benchmark/torchcor/volume.py).
    Kernel figures from HIP events on every 16th timed step."""
    import torch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    for N in sides:
        torch.cuda.reset_peak_memory_stats(dev)
        sim, el, its, t_setup = window(N, dev, warmup, steps, 16 if steps >= 256 else 1)
        n = sim.n_nodes
        kt = {k: round(ms / c * 1e3, 2) for k, (ms, c) in sim.kernel_times().items() if c}
        S = len(sim.ionic_model.STATES)
        ion_bytes = n * (8 + 8 * S + 8 + 8)
        B = alg_bytes(n, sim.system.nnz, S, column_bytes(sim.system),
                      xb=sim._X[0].element_size(), wb=sim._w.element_size())
        line = {"size_sweep": True, "side": N, "n_nodes": n, "nnz": sim.system.nnz,
                "steps": steps, "warmup": warmup, "t_window_ms": [warmup * 0.01, (warmup + steps) * 0.01],
                "ms_per_step": round(el / steps * 1e3, 4),
                "node_timesteps_per_s": n * steps / el,
                "pcg_iterations_per_step": sum(its) / steps + 1,
                "kernel_avg_us": kt,
                "kernel_alg_bytes": B,
                "ionic_GBps": round(ion_bytes / (kt["ionic"] * 1e-6) / 1e9, 1) if kt.get("ionic") else None,
                "setup_s": round(t_setup, 2),
                "peak_device_GB": round(torch.cuda.max_memory_allocated(dev) / 1e9, 2)}
        print(json.dumps(line), flush=True)
        sim.finish()
        del sim
        gc.collect()
        torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=101, help="nodes per cube side")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU work per storage type in the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-steady", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-shell", action="store_true")
    ap.add_argument("--served-requests", type=int, default=32)
    ap.add_argument("--shell-rcm", action="store_true",
                    help="also run the shell in plain reverse Cuthill-McKee order and in the "
                         "level order alone")
    ap.add_argument("--shell-steps", type=int, default=200)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--batch-warmup", type=int, default=200)
    ap.add_argument("--timing-period", type=int, default=16,
                    help="time the kernels of every k-th step with HIP events")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 5 plain device copies of this many MiB (the "
                         "calibration for HBM byte counters) and exit")
    ap.add_argument("--graph", type=int, default=0, choices=(0, 1),
                    help="1: every leg replays each step's fixed launches as a captured "
                         "HIP graph (timed steps launch eagerly)")
    ap.add_argument("--trace-summary", default="",
                    help="only summarise the last --steps steps of each pass in this "
                         "rocprofv3 kernel_trace.csv (no GPU needed) and exit")
    ap.add_argument("--size-sweep", default="",
                    help="only time the TTP cube at these comma-separated sides (e.g. "
                         "21,41,101,201), --warmup then --steps steps each, one JSON line "
                         "per size, and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the simulator runs on a single MI355X).", flush=True)
        sys.exit(2)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.trace_summary:
        trace_summary(args.trace_summary, args.steps)
        return
    if args.size_sweep:
        size_sweep([int(v) for v in args.size_sweep.split(",")], args.warmup, args.steps)
        return

    import torch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N = args.n
    # 1. the timed pass: events only when they leave >= 16 samples
    period = args.timing_period
    if args.no_kernel_timing or args.steps // max(period, 1) < 16:
        period = 0
    print(f"[bench] building {N}^3 workload; warmup {args.warmup}, timing {args.steps} steps",
          flush=True)
    graph = bool(args.graph)
    sim, elapsed, iters, t_setup = window(N, dev, args.warmup, args.steps, period, graph)
    n, nnz = sim.n_nodes, sim.system.nnz
    value = n * args.steps / elapsed
    S = len(sim.ionic_model.STATES)
    B = alg_bytes(n, nnz, S, column_bytes(sim.system), xb=sim._X[0].element_size(),
                  wb=sim._w.element_size())
    kt = sim.kernel_times() if period else {}
    n_elems = int(sim.elems.shape[0])
    from torchcor_amd import autotune
    tuning = {"table": str(autotune.TABLE_PATH.relative_to(ROOT)) if autotune._enabled else None,
              "pcg_blocks": sim.system.pcg_blocks,
              "pcg_slices_per_block": autotune.lookup("pcg_slices_per_block", n)}
    ops = host_operators(sim, N) if not args.no_cpu_baseline else None
    sim._free_sim()
    del sim
    torch.cuda.empty_cache()
    timing_note = f"timed pass, every {period}th step"
    if not period and not args.no_kernel_timing:
        # 2. a second pass over the same step window with every step timed
        print(f"[bench] kernel timing: the same {args.steps}-step window, every step", flush=True)
        sim, _, it2, _ = window(N, dev, args.warmup, args.steps, 1)
        kt = sim.kernel_times()
        timing_note = ("second pass over the same step window (same warm-up and steps), every "
                       f"step timed; its n_iter matched the timed pass: {it2 == iters}")
        sim._free_sim()
        del sim
        torch.cuda.empty_cache()
    kernels = kernel_stats(kt, B)
    roof = roofline(kernels, B, N) if kernels else None
    if roof is not None:
        # what a plain copy of the dominant kernel's bytes reaches from HBM
        cc = copy_ceiling(roof["alg_bytes"])
        roof["copy_ceiling_GBps"] = cc
        roof["frac_of_copy"] = round(roof["achieved"] / cc, 4)
    pcg = pcg_rate(kernels, B) if kernels else None
    # the same window with the step's fixed launches replayed as a HIP graph
    # (and eagerly when the timed pass used graphs), no kernel timing
    other = not graph
    print(f"[bench] the same window, graph replay {other}", flush=True)
    sim, el_o, it_o, _ = window(N, dev, args.warmup, args.steps, 0, other)
    graph_replay = {"ms_per_step": round((el_o if other else elapsed) / args.steps * 1e3, 4),
                    "eager_ms_per_step": round((elapsed if other else el_o) / args.steps * 1e3, 4),
                    "same_iterations": it_o == iters}
    sim._free_sim()
    del sim
    torch.cuda.empty_cache()
    # 3. a steady-state window (the wave well under way)
    steady = None
    if not args.no_steady:
        ws, ss = 200, 2000
        print(f"[bench] steady state: {ss} steps after {ws}", flush=True)
        sim, el, it3, _ = window(N, dev, ws, ss, 0 if args.no_kernel_timing else 16, graph)
        kst = kernel_stats(sim.kernel_times(), B) if not args.no_kernel_timing else {}
        steady = {"steps": ss, "t_window_ms": [ws * 0.01, (ws + ss) * 0.01],
                  "ms_per_step": round(el / ss * 1e3, 4), "node_timesteps_per_s": n * ss / el,
                  "pcg_iterations_per_step": sum(it3) / ss + 1,
                  "kernel_avg_us": {k: v["avg_us"] for k, v in kst.items()},
                  "roofline_frac": roofline(kst, B, N)["frac"] if kst else None,
                  "pcg_hbm_GBps": pcg_rate(kst, B) if kst else None}
        sim._free_sim()
        del sim
        torch.cuda.empty_cache()
    batched = None
    if not args.no_batched:
        batched = batched_config(dev, args.batch_steps, args.batch_warmup, graph)
        tuning["batch_max_blocks"] = batched["max_blocks"]
        gc.collect()
        torch.cuda.empty_cache()
    served = None
    if not args.no_served:
        served = served_config(dev, args.served_requests, graph=graph)
        tuning["serve_slices_per_block"] = sorted({autotune.lookup("serve_slices_per_block", m ** 3)
                                                   for m in range(11, 82)})
        gc.collect()
        torch.cuda.empty_cache()
    irregular = None
    if not args.no_shell:
        orders = (None, "rcm", "levels", "auto") if args.shell_rcm else (None, "auto")
        irregular = irregular_config(dev, args.shell_steps, 700, orders=orders)
        gc.collect()
        torch.cuda.empty_cache()
    cpu = None
    if ops is not None:
        cpu = cpu_baseline(ops, N, args.cpu_seconds)
    line = {
        "metric": METRIC, "value": value, "unit": "node-timesteps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"cube{N}_ttp_endo_f32: {n} nodes, {n_elems} P1 tets, "
                               f"nnz {nnz}, dt 0.01 ms, Jacobi-PCG tol 1e-5",
                   "n_nodes": n, "nnz": nnz, "t_window_ms": [args.warmup * 0.01,
                                                            (args.warmup + args.steps) * 0.01],
                   "pcg_iterations_per_step": sum(iters) / max(len(iters), 1) + 1},
        "pcg_hbm_GBps": pcg,
        "kernel_timing": timing_note,
        "kernels": kernels,
        "roofline": roof,
        "steady_state": steady,
        "tuning": tuning,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "irregular": irregular,
        "graph": graph,
        "graph_replay": graph_replay,
        "setup_s": round(t_setup, 2),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
