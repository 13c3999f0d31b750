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

Extra fields: per-kernel times measured with HIP events on the stream the
kernels run on, the PCG algorithmic HBM rate, the roofline of the dominant
kernel, and the CPU oracle timed on a bounded sample of the same workload.
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
    "ionic": 173_180_357,
    "rhs_init": 174_625_266,
    "pcg_spmv": 82_717_404,
    "pcg_update": 37_295_610,
}


def survey_electric_bytes(n, nnz):
    """Electric-phase algorithmic bytes of the unfused formulation with int32
    indices (B_rhs, B_init, B_iter): the definition behind 'PCG HBM GB/s'.
    The fused passes move fewer bytes than this (pcg_moved_GBps counts
    those), so the rate by this definition can exceed the HBM peak."""
    return (12 * nnz + 4 * (n + 1) + 12 * n, 8 * nnz + 4 * (n + 1) + 32 * n,
            8 * nnz + 4 * (n + 1) + 48 * n)


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


def cpu_baseline(N, mesh, steps):
    """The float64 CPU oracle on the same workload: `steps` time steps after its
    own (untimed) assembly, single-threaded numpy/scipy."""
    import numpy as np
    from oracle import ionic as oion
    from oracle import monodomain as omono
    from torchcor_amd import meshes
    nodes, elems, regions, fib = mesh
    m = omono.Monodomain(oion.TenTusscherPanfilov("ENDO", 0.01), T=100.0, dt=0.01)
    m.load_arrays(nodes, elems, regions, fib)
    m.add_conductivity([0], 0.5272, 0.2076)
    m.add_conductivity([1], 0.9074, 0.3332)
    m.add_stimulus(meshes.corner_nodes(N, 10), 0.0, 1.0, 100)
    print("[bench] cpu oracle: assembling", flush=True)
    m.assemble()
    m.assemble = lambda: None
    print(f"[bench] cpu oracle: {steps} steps", flush=True)
    t0 = time.perf_counter()
    with np.errstate(all="ignore"):
        m.solve(1e-5, 1e-5, 100, nsteps=steps)
    el = time.perf_counter() - t0
    return {"value": nodes.shape[0] * steps / el, "unit": "node-timesteps/s", "cores": 1,
            "kind": "port",
            "sample": f"{steps} time steps (t = 0.01..{steps * 0.01:.2f} ms) of the same "
                      f"{N}^3 TTP workload in float64 numpy/scipy, assembly excluded; "
                      f"{el:.1f} s"}


def batched_config(dev, steps, warmup):
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
    bat.start(1e-5, 1e-5, 100)
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
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
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
            "pcg_iterations_per_step_max": sum(bat.iterations[warmup:]) / steps + 1}


def irregular_config(dev, steps, warmup, n_target=640_000):
    """Irregular volume: the thick ellipsoidal shell (semi-axes 40/30/60 mm,
    wall 10 mm; jittered-lattice Delaunay, about 640k nodes, randomly
    permuted node numbers), four conductivity regions, transmural fibres,
    five stimulus patches (0 and 5 ms onsets), TTP ENDO float32.  Runs the
    same steps in the mesh's own (scattered) order, in reverse Cuthill-McKee
    order and in the simulator's default locality order (Cuthill-McKee levels
    along a Z-order curve)."""
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
    for reorder in (None, "rcm", "levels"):
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
                    "kernel_avg_us": kt,
                    "pcg_iterations_per_step": sum(sim.iterations[warmup:]) / steps + 1}
        sim.finish()
        del sim
    out["speedup_rcm"] = round(out["mesh_order"]["ms_per_step"] / out["rcm"]["ms_per_step"], 2)
    out["speedup_levels"] = round(out["mesh_order"]["ms_per_step"] / out["levels"]["ms_per_step"], 2)
    return out


def served_config(dev, n_req, budget_nodes=4_000_000):
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
                      device=dev)
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
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
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
            "sims_per_hour": round(summary["sims_per_hour"], 1),
            "latency_per_sim_s_p50": round(summary["latency_per_sim_s_p50"], 3),
            "latency_per_sim_s_p99": round(summary["latency_per_sim_s_p99"], 3),
            "node_timesteps": int(sum(r.nodes.shape[0] * int(r.T / 0.01) for r in reqs)),
            "back_to_back_single_runs": b2b,
            "speedup_vs_back_to_back": round(summary["sims_per_hour"] / b2b["sims_per_hour"], 2)}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=101, help="nodes per cube side")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-requests", type=int, default=32)
    ap.add_argument("--shell", action="store_true",
                    help="add the irregular ellipsoidal-shell workload (mesh order vs RCM)")
    ap.add_argument("--shell-steps", type=int, default=1000)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--batch-warmup", type=int, default=200)
    ap.add_argument("--timing-period", type=int, default=16,
                    help="time the kernels of every k-th step with HIP events")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 5 plain device copies of this many MiB (the "
                         "calibration for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the simulator runs on a single MI355X).", flush=True)
        sys.exit(2)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return

    import torch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N = args.n
    print(f"[bench] building {N}^3 workload", flush=True)
    sim, mesh = build_sim(N, dev)
    t0 = time.perf_counter()
    sim.start(1e-5, 1e-5, 100, calculate_AT_RT=True, linear_guess=True,
              snapshot_interval=1e9, verbose=False)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    n, nnz = sim.n_nodes, sim.system.nnz
    print(f"[bench] assembled n={n} nnz={nnz} in {t_setup:.2f} s; warmup {args.warmup}", flush=True)
    sim.advance(args.warmup)
    torch.cuda.synchronize()
    if not args.no_kernel_timing:
        sim.set_kernel_timing(True, period=args.timing_period)
    it0 = len(sim.iterations)
    print(f"[bench] timing {args.steps} steps", flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.advance(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    iters = sim.iterations[it0:]
    value = n * args.steps / elapsed
    kt = {} if args.no_kernel_timing else sim.kernel_times()
    S = len(sim.ionic_model.STATES)
    B = alg_bytes(n, nnz, S, column_bytes(sim.system), xb=sim._X[0].element_size(),
                  wb=sim._w.element_size())
    kernels = {}
    for k, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        avg_s = ms / cnt / 1e3
        kernels[k] = {"launches": cnt, "avg_us": round(avg_s * 1e6, 2), "total_ms": round(ms, 2)}
        if k in B:
            kernels[k].update(alg_bytes=B[k], GBps=round(B[k] / avg_s / 1e9, 1))
    roof = None
    pcg_rate = pcg_moved = None
    if kernels:
        dom = max((k for k in kernels if k in B), key=lambda k: kernels[k]["total_ms"])
        a = kernels[dom]["GBps"]
        traffic = MEASURED_TRAFFIC_101.get(dom) if N == 101 else None
        roof = {"kernel": dom, "bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(a / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "alg_bytes": B[dom]}
        b_rhs, b_init, b_iter = survey_electric_bytes(n, nnz)
        n_it = kernels.get("pcg_spmv", {}).get("launches", 0)
        n_st = kernels.get("rhs_init", {}).get("launches", 0)
        ele_bytes = n_st * (b_rhs + b_init) + n_it * b_iter
        ele_s = sum(kernels[k]["total_ms"] for k in ("rhs_init", "pcg_spmv", "pcg_update",
                                                     "pcg_check") if k in kernels) / 1e3
        pcg_rate = round(ele_bytes / ele_s / 1e9, 1)
        # the same phase by the bytes the fused passes actually move
        moved = sum(kernels[k]["launches"] * B[k] for k in ("rhs_init", "pcg_spmv", "pcg_update")
                    if k in kernels)
        pcg_moved = round(moved / ele_s / 1e9, 1)
    # release the single-run workload's device memory before the next parts,
    # so that each part runs on a device holding only its own data
    n_elems = int(sim.elems.shape[0])
    sim._free_sim()
    del sim
    torch.cuda.empty_cache()
    batched = None
    if not args.no_batched:
        batched = batched_config(dev, args.batch_steps, args.batch_warmup)
        gc.collect()
        torch.cuda.empty_cache()
    served = None
    if not args.no_served:
        served = served_config(dev, args.served_requests)
        gc.collect()
        torch.cuda.empty_cache()
    irregular = None
    if args.shell:
        irregular = irregular_config(dev, args.shell_steps, 700)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(N, mesh, args.cpu_steps)
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
        "pcg_hbm_GBps": pcg_rate,
        "pcg_moved_GBps": pcg_moved,
        "kernels": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "irregular": irregular,
        "setup_s": round(t_setup, 2),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
