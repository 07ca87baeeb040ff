"""Benchmark: photon paths per second on one MI355X, single ALB shot.

One step = one full pass of the hot path over one shot: forward trace of
N_f photons (map records into HBM), four photon-map builds, backward trace of
N_b sensor paths, k-NN gather into FP64 waveforms.  Default workload: the
400 m / 5 m scene at N_f = 1e8, N_b = 2.5e7 (k = 100).  value = (N_f + N_b) /
(wall seconds per step), wall time bracketed by device synchronisation.

Also reported: per-phase times (HIP events on the launch stream), the
roofline position of the dominant kernel (k-NN: algorithmic bytes and flops
from the kernel's own event counters over its launch time; the binding unit
from its hardware counters in profiles/knn_kernel_counters.json; HBM traffic
per launch from profiles/knn_kernel_traffic.json, rocprofv3 FETCH_SIZE /
WRITE_SIZE runs scaled by a copy kernel's reading), and the CPU baseline: the
same scene at 1e5 + 2.5e4 photons in full (time-stepped C oracle trace on
this host's cores, scipy cKDTree k-NN, oracle gather), with the GPU's rate on
that same small workload beside it.

Radius mode (the 1e8 shot): the fixed-radius gather of a 256 x 256 footprint
grid on the seafloor plane (1 cm pitch, r = 0.05 m) over the BOTTOM and
SCATTER maps of one more 1e8-photon forward pass; per map the gather time
(HIP events), records taken and examined per query, and algorithmic GB/s.

The other forms of the metric:
  * batched: the first --sweep-full scenes of the 1024-scene Latin-hypercube
    sweep (seed 1) at their full photon counts 10^U[5,7], batched in
    memory-bounded scene groups vs plain per-scene runs; batched_small: the
    first --sweep-scenes scenes at 1/100 of those counts (the regime where
    per-scene launch and synchronisation costs dominate);
  * adjoint: --adjoint-scenes sweep scenes at --adjoint-photons forward (1/4
    backward) photons, MSE against targets from perturbed parameters: wall
    time of forward + path-replay adjoint over the forward alone, in
    memory-bounded scene groups;
  * adjoint_vs_differences (opt-in, --adjoint-fd-replicas R): the adjoint's
    a, b, g gradients against central differences with common random
    numbers, paired per replica over R independent replicas of each of
    --adjoint-fd-scenes sweep scenes (MSE or linearised loss);
  * inversion: Adam over the batched adjoint on --inversion-scenes synthetic
    scenes (unknown depth, b, albedo), --inversion-iters iterations; depth
    RMSE and seconds per iteration, batched vs plain; with
    --inversion-full-iteration also one iteration over all 4096 scenes.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--forward N]
       [--backward N] [--no-cpu] [--no-batched] [--no-adjoint] [--no-radius]
       [--no-inversion] [--sweep-full N] [--adjoint-scenes N]
       [--inversion-scenes N] [--inversion-full-iteration]
       [--adjoint-fd-replicas R --adjoint-fd-scenes N --adjoint-fd-first S
        --adjoint-fd-photons N --adjoint-fd-eps E --adjoint-fd-loss mse|linear
        --adjoint-fd-params abg]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "photon paths/s on 1 GPU, plain vs batched; adjoint/forward time ratio"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# algorithmic work per event (flops: add/sub/mul/div/cmp = 1, FMA = 2,
# sqrt/log/exp/sin/cos/acos/pow = 1)
F_FWD_EMIT, F_BWD_EMIT, F_AIR, F_WATER_STEP = 30, 50, 12, 24
F_DEPTH_SOLVE = 12  # one closed-form optical-depth jump (or single step)
F_SCATTER, F_ENTER, F_EXIT, F_FLOOR = 95, 70, 80, 70
M_TYPE = (0, 30, 85, 0)  # per-neighbour multiplier flops by photon type


def trace_flops(ctr, forward: bool) -> float:
    from pyalbsim_amd import abi

    # the water sub-steps are reproduced in closed form: count the jumps, not the sub-steps
    return float((F_FWD_EMIT if forward else F_BWD_EMIT) * ctr[abi.CTR_PATHS] + F_AIR * ctr[abi.CTR_AIR_LEGS]
                 + F_DEPTH_SOLVE * ctr[abi.CTR_DEPTH_SOLVES] + F_SCATTER * ctr[abi.CTR_SCATTER]
                 + F_ENTER * ctr[abi.CTR_ENTER] + F_EXIT * ctr[abi.CTR_EXIT] + F_FLOOR * ctr[abi.CTR_FLOOR])


def host_cpu() -> dict:
    """The host's CPU model, logical CPUs (os.cpu_count) and physical cores."""
    model, phys = None, set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if "physical id" in cur and "core id" in cur:
                        phys.add((cur["physical id"], cur["core id"]))
                    cur = {}
                    continue
                k_, v = (x.strip() for x in line.split(":", 1))
                cur[k_] = v
                if k_ == "model name" and model is None:
                    model = v
        if "physical id" in cur and "core id" in cur:
            phys.add((cur["physical id"], cur["core id"]))
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(phys) or None}


def cpu_baseline(model, key, n_f: int, n_b: int, threads: int):
    """The C oracle on n_f + n_b paths of the scene, on `threads` host threads:
    the time-stepped trace of both passes (each thread a range of photons),
    the maps per type, the k-NN with scipy's cKDTree (the k-d tree query the
    reference's gather uses, photon_map_index.py:17, core/simulation.py:
    139-142) on all threads, and the oracle's gather over those neighbour
    sets.  Returns (paths/s, wall seconds, seconds per stage)."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from scipy.spatial import cKDTree

    import oracle
    from pyalbsim_amd import abi

    oracle.lib()
    stages = {}
    t0 = time.perf_counter()

    def chunks(n, parts):
        step = max(1, (n + parts - 1) // parts)
        return [(b, min(step, n - b)) for b in range(0, n, step)]

    with ThreadPoolExecutor(threads) as ex:
        # photon ranges several times the thread count, so slow paths even out
        fw = list(ex.map(lambda bc: oracle.trace(model, key, abi.PASS_FORWARD, bc[0], bc[1])[0],
                         chunks(n_f, 8 * threads)))
        t1 = time.perf_counter()
        bw = list(ex.map(lambda bc: oracle.trace(model, key, abi.PASS_BACKWARD, bc[0], bc[1])[0],
                         chunks(n_b, 8 * threads)))
        t2 = time.perf_counter()
        maps = [{f: np.concatenate([c[t][f] for c in fw]) for f in fw[0][t]} for t in range(4)]
        qs = [{f: np.concatenate([c[t][f] for c in bw]) for f in bw[0][t]} for t in range(4)]
        t_knn = 0.0
        jobs = []
        for t in range(4):
            nq = len(qs[t]["x"])
            if nq == 0 or len(maps[t]["x"]) < model.k:
                continue
            ta = time.perf_counter()
            pts = np.stack([maps[t]["x"], maps[t]["y"], maps[t]["z"]], 1).astype(np.float64)
            qp = np.stack([qs[t]["x"], qs[t]["y"], qs[t]["z"]], 1)
            _, nb = cKDTree(pts).query(qp, k=model.k, workers=threads)
            t_knn += time.perf_counter() - ta
            for b, c in chunks(nq, 8 * threads):
                jobs.append((t, {f: v[b:b + c] for f, v in qs[t].items()}, nb[b:b + c]))
        t3 = time.perf_counter()
        list(ex.map(lambda j: oracle.gather(model, j[0], maps[j[0]], j[1], nbrs=j[2]), jobs))
    wall = time.perf_counter() - t0
    stages = {"forward_trace_s": t1 - t0, "backward_trace_s": t2 - t1, "knn_s": t_knn, "gather_s": time.perf_counter() - t3}
    return (n_f + n_b) / wall, wall, stages


def gpu_small_shot(model, key, n_f: int, n_b: int, repeats: int = 5):
    """GPU paths/s of the plain form on a small shot (n_f + n_b photons), the
    median of `repeats` timed runs after one warm-up."""
    import numpy as np
    import torch

    from pyalbsim_amd.execution.linear import simulate

    simulate(model, n_f, n_b, key)
    torch.cuda.synchronize()
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        simulate(model, n_f, n_b, key)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return (n_f + n_b) / float(np.median(ts)), float(np.median(ts))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--forward", type=float, default=1e8)
    ap.add_argument("--backward", type=float, default=2.5e7)
    ap.add_argument("--cpu-forward", type=int, default=100000)
    ap.add_argument("--cpu-backward", type=int, default=25000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-radius", action="store_true")
    ap.add_argument("--no-inversion", action="store_true")
    ap.add_argument("--inversion-scenes", type=int, default=512)
    ap.add_argument("--inversion-iters", type=int, default=10)
    ap.add_argument("--inversion-lr", type=float, default=0.02)
    ap.add_argument("--radius-grid", type=int, default=256)
    ap.add_argument("--sweep-scenes", type=int, default=256)
    ap.add_argument("--sweep-full", type=int, default=32,
                    help="sweep scenes run at their full photon counts (grouped batched vs plain); 0: none")
    ap.add_argument("--adjoint-scenes", type=int, default=256)
    ap.add_argument("--adjoint-photons", type=float, default=1e6)
    ap.add_argument("--adjoint-fd-replicas", type=int, default=0,
                    help="also compare the adjoint's a, b, g gradients with central CRN differences over this "
                         "many independent replicas of each of the first --adjoint-fd-scenes sweep scenes")
    ap.add_argument("--adjoint-fd-scenes", type=int, default=2)
    ap.add_argument("--adjoint-fd-first", type=int, default=0, help="first sweep scene of that comparison")
    ap.add_argument("--adjoint-fd-photons", type=float, default=1e5, help="forward photons per replica")
    ap.add_argument("--adjoint-fd-eps", type=float, default=5e-2, help="relative step of the differences")
    ap.add_argument("--adjoint-fd-params", default="abg", help="parameters compared, of a, b, g")
    ap.add_argument("--adjoint-fd-loss", choices=("mse", "linear"), default="mse",
                    help="mse: sum (W - T)^2; linear: sum 2 (W_ref - T) W with W_ref the scene's mean waveform "
                         "from separate runs (the MSE linearised at the mean: no variance term)")
    ap.add_argument("--inversion-full-iteration", action="store_true",
                    help="also time one inversion iteration over 4096 scenes at 2e4 photons per scene "
                         "(and scale it to 2e5 photons x 200 iterations)")
    ap.add_argument("--inversion-full-photons", type=float, default=2e4,
                    help="forward photons per scene of --inversion-full-iteration (2e5: the workload's count)")
    ap.add_argument("--inversion-full-group", type=int, default=256)
    ap.add_argument("--inversion-full-iters", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 2 GiB device copy at the end (reference reading for HBM byte counters)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from pyalbsim_amd import abi, runtime
    from pyalbsim_amd.execution.linear import simulate
    from pyalbsim_amd.model import build_scene
    from pyalbsim_amd.sweep import shot_config

    runtime.load_library()
    model = build_scene(shot_config(depth=5.0, a=0.1, b=0.2, albedo=0.15, roughness=0.05))
    n_f, n_b = int(args.forward), int(args.backward)
    ds = runtime.DeviceScene(model)
    key = (0x5EED, 0)

    def step(events=None, kernel_events=None):
        return simulate(model, n_f, n_b, key, dscene=ds, events=events, kernel_events=kernel_events)

    for i in range(args.warmup):
        step()
        print(f"warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    events = {}
    kev = {}
    t0 = time.perf_counter()
    res = None
    for i in range(args.steps):
        res = step(events, kev)
        print(f"step {i + 1}/{args.steps} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    value = (n_f + n_b) / (wall / args.steps)

    phases = {}
    names = ["start", "forward", "build", "backward", "gather"]
    for a, b in zip(names[:-1], names[1:]):
        phases[b] = sum(s.elapsed_time(e) for s, e in zip(events[a], events[b])) / args.steps
    cf, cb = res.counters_fwd, res.counters_bwd
    flops_fwd = trace_flops(cf, True)
    flops_bwd = trace_flops(cb, False)
    nq = float(cb[abi.CTR_QUERIES])
    cand = float(cb[abi.CTR_CANDIDATES])
    nodes = float(cb[abi.CTR_NODES])
    k = model.k
    qc = res.query_counts
    g_flops = sum(qc[t] * (4 + (7 + M_TYPE[t]) * k) for t in range(4)) + 15 * cand
    # k-NN algorithmic work per query (the gather's counting rules): bytes
    # 12 per candidate's coordinates + 20 per neighbour (index, direction,
    # energy, time) + 8 per tree node visited; flops 4 + 15 per candidate +
    # 7 per neighbour
    knn_bytes = 12 * cand + 20 * k * nq + 8 * nodes
    knn_flops = 4 * nq + 15 * cand + 7 * k * nq
    # the same with this tree's node size and the neighbour payload left to
    # the contribution kernel: 12 per candidate + 4 per neighbour index
    # written + 64 per node line
    knn_bytes_lines = 12 * cand + 4 * k * nq + 64 * nodes
    knn_ms = [s_.elapsed_time(e_) for s_, e_ in kev["knn"]]
    contrib_ms = [s_.elapsed_time(e_) for s_, e_ in kev["contrib"]]
    knn_launches_per_step = len(knn_ms) / args.steps
    knn_avg_ms = sum(knn_ms) / len(knn_ms)
    t_launch = knn_avg_ms / 1e3
    per_launch = 1.0 / knn_launches_per_step
    roof = {"kernel": "knn_kernel", "bound": "valu",
            "achieved": knn_bytes * per_launch / t_launch / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "byte_model": "12 B per candidate + 20 B per neighbour + 8 B per tree node, per query",
            "avg_launch_ms": knn_avg_ms, "launches_per_step": knn_launches_per_step,
            "achieved_node_lines": knn_bytes_lines * per_launch / t_launch / 1e9,
            "flops": {"achieved": knn_flops * per_launch / t_launch / 1e12, "peak": PEAK_FP64_TFLOPS,
                      "unit": "TFLOP/s"}}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["frac_node_lines"] = roof["achieved_node_lines"] / roof["peak"]
    roof["flops"]["frac"] = roof["flops"]["achieved"] / roof["flops"]["peak"]
    roof["traffic"] = None
    tpath = os.path.join(ROOT, "profiles", "knn_kernel_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        roof["traffic"] = tr.get("bytes_per_launch")
        roof["traffic_source"] = tr.get("source")
    cpath = os.path.join(ROOT, "profiles", "knn_kernel_counters.json")
    if os.path.exists(cpath):
        # the binding unit from the kernel's counters: VALU issue against DRAM traffic
        with open(cpath) as f:
            cn = json.load(f)
        roof["counters"] = {x: cn.get(x) for x in ("valu_issue_frac", "salu_issue_frac", "lds_bank_conflict_frac",
                                                   "dram_frac", "source")}
        roof["bound"] = "valu" if (cn.get("valu_issue_frac") or 0) >= (cn.get("dram_frac") or 0) else "hbm"
    # whole gather (k-NN + contributions): node lines, candidates, neighbour
    # indices, and per (query, neighbour) pair the record payload and row data
    g_bytes = knn_bytes_lines + (20 + 48) * k * nq
    t_gather = phases["gather"] / 1e3
    t_trace = (phases["forward"] + phases["backward"]) / 1e3

    out = {
        "metric": METRIC, "value": value, "unit": "photon paths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"single ALB shot (400 m, nadir, 1 mrad, FOV 20 mrad, 5 m water, a=0.1, b=0.2, "
                               f"albedo 0.15, 0.1 ns bins), plain form, {n_f:.3g} forward + {n_b:.3g} "
                               f"backward photons, k=100",
                   "forward_photons": n_f, "backward_photons": n_b, "k": k,
                   "map_records": res.map_counts, "gather_queries": qc},
        "phases_ms": phases,
        "forward_paths_per_s": n_f / (phases["forward"] / 1e3) if phases["forward"] > 0 else None,
        "backward_paths_per_s": n_b / (phases["backward"] / 1e3) if phases["backward"] > 0 else None,
        "kernel_ms_per_step": {"knn": sum(knn_ms) / args.steps, "contrib": sum(contrib_ms) / args.steps},
        "roofline": roof,
        "trace_tflops": (flops_fwd + flops_bwd) / t_trace / 1e12,
        "gather_gbs": g_bytes / t_gather / 1e9 if t_gather > 0 else None,
        "gather_candidates_per_query": cand / max(nq, 1.0),
        "gather_nodes_per_query": nodes / max(nq, 1.0),
        "knn_steps_per_query": float(cb[abi.CTR_KNN_STEPS]) / max(nq, 1.0),
        "knn_sorts_per_query": float(cb[abi.CTR_KNN_SORTS]) / max(nq, 1.0),
        "knn_descent_per_query": float(cb[abi.CTR_KNN_DESCENT]) / max(nq, 1.0),
        "water_steps_per_path": float(cf[abi.CTR_WATER_STEPS]) / max(float(cf[abi.CTR_PATHS]), 1.0),
        "depth_solves_per_path": float(cf[abi.CTR_DEPTH_SOLVES]) / max(float(cf[abi.CTR_PATHS]), 1.0),
        "cpu_baseline": None,
    }
    del res
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        rate, cwall, stages = cpu_baseline(model, key, args.cpu_forward, args.cpu_backward, threads)
        g_rate, g_s = gpu_small_shot(model, key, args.cpu_forward, args.cpu_backward)
        out["cpu_baseline"] = {"value": rate, "unit": "photon paths/s", "cores": threads, "kind": "port",
                               "host": host_cpu(), "stages": stages,
                               "sample": f"the same scene at {args.cpu_forward} forward + {args.cpu_backward} "
                                         f"backward photons, in full: time-stepped C oracle trace on {threads} "
                                         f"threads, scipy cKDTree k-NN, oracle gather; {cwall:.1f} s wall",
                               "gpu_same_workload": {"value": g_rate, "unit": "photon paths/s", "seconds": g_s},
                               "gpu_over_cpu": g_rate / rate}
        print("cpu baseline done", file=sys.stderr, flush=True)
    def leg(name, fn, *a, **kw):
        # each finished leg also goes to stderr, so a later failure keeps it
        out[name] = fn(*a, **kw)
        print(json.dumps({name: out[name]}), file=sys.stderr, flush=True)

    if not args.no_radius:
        leg("radius_mode", bench_radius, model, ds, key, n_f, args.radius_grid)
    if not args.no_batched:
        if args.sweep_full > 0:
            leg("batched", bench_batched_full, args.sweep_full)
        if args.sweep_scenes > 0:
            leg("batched_small", bench_batched, args.sweep_scenes)
    if not args.no_adjoint:
        leg("adjoint", bench_adjoint, args.adjoint_scenes, int(args.adjoint_photons))
        if args.adjoint_fd_replicas > 0:
            leg("adjoint_vs_differences", bench_adjoint_fd, args.adjoint_fd_scenes, args.adjoint_fd_replicas,
                first=args.adjoint_fd_first, loss=args.adjoint_fd_loss, nf=int(args.adjoint_fd_photons),
                group=max(1, int(128 * 1e5 // args.adjoint_fd_photons)), eps=args.adjoint_fd_eps,
                params=args.adjoint_fd_params)
    if not args.no_inversion:
        leg("inversion", bench_inversion, args.inversion_scenes, args.inversion_iters, lr=args.inversion_lr)
        if args.inversion_full_iteration:
            leg("inversion_full_iteration", bench_inversion_full_iteration, n_forward=int(args.inversion_full_photons),
                group=args.inversion_full_group, iterations=args.inversion_full_iters)
    if args.copy_calibration:
        x = torch.empty(1 << 28, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = {"read": x.numel() * 8, "write": y.numel() * 8}
    print(json.dumps(out), flush=True)
    return 0


def bench_radius(model, ds, key, n_f: int, grid: int, radius: float = 0.05, pitch: float = 0.01,
                 repeats: int = 3):
    """Fixed-radius gather of a footprint grid over the BOTTOM and SCATTER maps."""
    import torch

    from pyalbsim_amd import abi, runtime
    from pyalbsim_amd.footprint import footprint_queries, upload_queries

    recs, rc = runtime.trace(ds, key, abi.PASS_FORWARD, n_f)
    types = (abi.BOTTOM_REFLECTION, abi.SCATTER)
    maps = {t: runtime.build_map(recs[t], rc[t], runtime.MAP_DIMS[t]) for t in types}
    del recs
    sink = upload_queries(footprint_queries(model, grid, pitch))
    nq = grid * grid
    out = {"workload": f"{grid}x{grid} footprint grid on the seafloor plane around the beam's floor point, "
                       f"{pitch * 100:g} cm pitch, r = {radius} m, over the maps of a {n_f:.3g}-photon forward pass",
           "queries": nq, "radius_m": radius, "types": {}}
    for t in types:
        wf = torch.zeros(model.n_bins, dtype=torch.float64, device="cuda")
        runtime.radius_gather(ds, maps[t], t, sink, nq, radius, wf, model.gate())  # warm-up
        ctr = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device="cuda")
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(repeats):
            wf.zero_()
            runtime.radius_gather(ds, maps[t], t, sink, nq, radius, wf, model.gate(), counters=ctr, check=False)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / repeats
        c = ctr.cpu().numpy().view("uint64").astype(float) / repeats
        if c[abi.CTR_OVERFLOW]:
            raise runtime.AlbError("radius gather stack overflow")
        taken, cand, nodes = c[abi.CTR_NEIGHBOURS], c[abi.CTR_CANDIDATES], c[abi.CTR_NODES]
        # bytes: coordinates of examined records (12 B), payload of taken ones
        # (dir 12 B for SCATTER, energy 4 B, time 4 B), node lines (64 B)
        pay = 20 if t == abi.SCATTER else 8
        gbytes = 12 * cand + pay * taken + 64 * nodes
        out["types"][abi.TYPE_NAMES[t]] = {
            "map_records": int(rc[t]), "ms": ms, "queries_per_s": nq / (ms / 1e3),
            "taken_per_query": taken / nq, "examined_per_query": cand / nq, "nodes_per_query": nodes / nq,
            "records_per_s": taken / (ms / 1e3), "algorithmic_gbs": gbytes / (ms / 1e3) / 1e9,
            "waveform_sum": float(wf.sum().item())}
    del maps
    print("radius mode done", file=sys.stderr, flush=True)
    return out


def bench_inversion(n_scenes: int, iters: int, n_forward: int = 20000, lr: float = 0.02, plain_scenes: int = 32):
    """Adam over the batched adjoint on synthetic inversion scenes: depth RMSE
    and seconds per iteration, batched form vs plain (one scene per launch;
    timed on the first plain_scenes scenes and scaled to n_scenes)."""
    import numpy as np

    from pyalbsim_amd.inversion import invert, inversion_scenes, inversion_tilts

    th = inversion_scenes(n_scenes, seed=3)
    tilts = inversion_tilts(n_scenes, seed=3)
    rb = invert(th, n_forward, n_forward // 4, iterations=iters, lr=lr, tilts=tilts, group_scenes=256,
                log=(lambda m: print(m, file=sys.stderr, flush=True)) if iters > 20 else None)
    print("inversion batched done", file=sys.stderr, flush=True)
    ps = min(plain_scenes, n_scenes)
    rp = invert(th[:ps], n_forward, n_forward // 4, iterations=2, lr=lr, form="plain", tilts=tilts[:ps])
    print("inversion plain done", file=sys.stderr, flush=True)
    sb = float(np.median(rb.seconds_per_iteration[1:] or rb.seconds_per_iteration))
    sp = float(np.median(rp.seconds_per_iteration)) * n_scenes / ps
    return {"workload": f"{n_scenes} scenes, depth U[2,40] m, b U[0.1,1.5] /m, albedo U[0.05,0.4] unknown "
                        f"(a = 0.1 /m, g = 0.9 known; roughness from wind U[0,10] m/s; floor tilted U[0,15] deg, "
                        f"azimuth U[0,360) deg, known), "
                        f"{n_forward} forward + {n_forward // 4} backward photons per scene per pass, "
                        f"Adam lr {lr:g} on [0,1]-scaled unknowns, depth started from the target's "
                        f"floor-return leading edge (mid-range where none shows), {iters} iterations; "
                        f"batched in groups of 256 scenes; plain form timed over 2 iterations of the first {ps} "
                        f"scenes, scaled to {n_scenes}",
            "depth_rmse_start_m": rb.depth_rmse[0], "depth_rmse_end_m": rb.depth_rmse[-1],
            "depth_rmse_every_10": rb.depth_rmse[::10],
            "loss_start": rb.loss[0], "loss_end": rb.loss[-1],
            "batched_s_per_iteration": sb, "plain_s_per_iteration": sp, "speedup": sp / sb}


def bench_inversion_full_iteration(n_scenes: int = 4096, n_forward: int = 20000, group: int = 256,
                                   iterations: int = 2):
    """Adam iterations over every inversion scene (batched, in groups of
    `group` scenes); seconds of the last iteration at n_forward photons and
    the estimate for 2e5 photons (the work is linear in photons) and 200
    iterations."""
    from pyalbsim_amd.inversion import invert, inversion_scenes, inversion_tilts

    th = inversion_scenes(n_scenes, seed=3)
    tilts = inversion_tilts(n_scenes, seed=3)
    r = invert(th, n_forward, n_forward // 4, iterations=iterations, lr=1e-2, tilts=tilts, group_scenes=group,
               graphs=iterations > 1, log=lambda m: print(m, file=sys.stderr, flush=True))
    s_it = float(r.seconds_per_iteration[-1])
    est = s_it * (2e5 / n_forward)
    print("inversion full iteration done", file=sys.stderr, flush=True)
    timed = "the second timed" if iterations > 1 else "timed; every pass eager, no HIP graphs"
    return {"workload": f"{n_scenes} inversion scenes, {n_forward} forward + {n_forward // 4} backward photons per "
                        f"scene, {iterations} Adam iteration(s) ({timed}) in groups of {group} scenes",
            "s_per_iteration": s_it, "estimated_s_per_iteration_at_2e5_photons": est,
            "estimated_s_for_200_iterations_at_2e5_photons": 200 * est,
            "depth_rmse": r.depth_rmse}


def bench_batched(n_scenes: int):
    """Plain per-scene runs vs one batched run over the same sweep scenes at
    1/100 of their photon counts."""
    import torch

    from pyalbsim_amd.execution.batched import simulate_batched
    from pyalbsim_amd.execution.linear import simulate
    from pyalbsim_amd.sweep import hypercube

    sc = hypercube(1024, 1, (3.0, 5.0))[:n_scenes]
    models = [x.model for x in sc]
    nf = [x.n_forward for x in sc]
    nb = [x.n_backward for x in sc]
    paths = sum(nf) + sum(nb)
    key = (0xBA7C, 1)
    from pyalbsim_amd.autotune import tune_batched

    t0 = time.perf_counter()
    block, refill, _ = tune_batched(models, nf, key)  # launch parameters for this sweep (cached)
    t_tune = time.perf_counter() - t0
    simulate_batched(models[:8], nf[:8], nb[:8], key, strict=False)  # warm-up
    simulate(models[0], nf[0], nb[0], key, strict=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = simulate_batched(models, nf, nb, key, strict=False, block=block, refill=refill)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    counts = (res.map_counts, res.query_counts)
    del res
    print(f"batched {n_scenes} scenes done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for s in range(n_scenes):
        simulate(models[s], nf[s], nb[s], key, strict=False, scene_id=s)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    print(f"plain {n_scenes} scenes done", file=sys.stderr, flush=True)
    return {"workload": f"first {n_scenes} of 1024 Latin-hypercube sweep scenes (seed 1; depth U[1,30] m, "
                        f"a U[0.02,0.5], b U[0.05,2], HG g U[0.8,0.95], albedo U[0.05,0.4], wind U[0,10] m/s), "
                        f"photons 10^U[3,5] per scene (1/100 of the sweep's 10^U[5,7]), backward = forward / 4",
            "forward_photons": int(sum(nf)), "backward_photons": int(sum(nb)),
            "map_records": counts[0], "gather_queries": counts[1],
            "batched_paths_per_s": paths / tb, "plain_paths_per_s": paths / tp, "batched_s": tb, "plain_s": tp,
            "speedup": tp / tb, "autotuned": {"block": block, "refill": refill, "tuning_s": t_tune}}


def bench_batched_full(n_scenes: int):
    """The first n_scenes sweep scenes at their full photon counts (10^U[5,7]):
    the grouped batched form (memory-bounded scene groups) vs plain per-scene
    runs."""
    import torch

    from pyalbsim_amd.autotune import tune_batched
    from pyalbsim_amd.execution.linear import simulate
    from pyalbsim_amd.sweep import hypercube, run_sweep

    sc = hypercube(1024, 1, (5.0, 7.0))[:n_scenes]
    nf = [x.n_forward for x in sc]
    nb = [x.n_backward for x in sc]
    paths = sum(nf) + sum(nb)
    key = (0xBA7C, 1)
    # the tuning pilot: a sixteenth of each scene's photons, at most 5e6 in all
    block, refill, _ = tune_batched([x.model for x in sc], nf, key, pilot_fraction=min(1 / 16, 5e6 / sum(nf)))
    torch.cuda.synchronize()
    res = run_sweep(sc, key, block=block, refill=refill, strict=False,
                    log=lambda m: print(m, file=sys.stderr, flush=True))
    tb = res.seconds
    print(f"batched full-size {n_scenes} scenes done ({len(res.groups)} groups)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for s in range(n_scenes):
        simulate(sc[s].model, nf[s], nb[s], key, strict=False, scene_id=s)
        if s % 8 == 7:
            print(f"plain full-size scene {s + 1}/{n_scenes}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    print(f"plain full-size {n_scenes} scenes done", file=sys.stderr, flush=True)
    return {"workload": f"first {n_scenes} of 1024 Latin-hypercube sweep scenes (seed 1) at full photon counts "
                        f"10^U[5,7], backward = forward / 4; batched in memory-bounded scene groups",
            "forward_photons": int(sum(nf)), "backward_photons": int(sum(nb)), "groups": len(res.groups),
            "batched_paths_per_s": paths / tb, "plain_paths_per_s": paths / tp, "batched_s": tb, "plain_s": tp,
            "speedup": tp / tb, "autotuned": {"block": block, "refill": refill}}


def bench_adjoint(n_scenes: int, photons: int = 1000000):
    """Forward vs forward + path-replay adjoint on sweep scenes (photons
    forward, photons / 4 backward per scene), MSE loss against targets from
    parameters perturbed by U[-10 %, 10 %] (seed 2).  The scenes run in
    memory-bounded groups (the adjoint keeps the sinks, maps and neighbour
    rows of its group); times are summed over the groups."""
    import numpy as np
    import torch

    from pyalbsim_amd import adjoint
    from pyalbsim_amd.execution.batched import simulate_batched
    from pyalbsim_amd.sweep import estimate_group_bytes, free_device_bytes, hypercube, plan_groups

    sc = hypercube(256, 1)[:n_scenes]
    th = np.array([[x.params[k] for k in adjoint.PARAMS] for x in sc])
    pert = th * (1 + np.random.default_rng(2).uniform(-0.1, 0.1, th.shape))
    pert[:, 2] = np.clip(pert[:, 2], 0.5, 0.97)
    models = [adjoint.model_from_theta(t) for t in th]
    tmodels = [adjoint.model_from_theta(t) for t in pert]
    nf, nb = int(photons), int(photons) // 4
    key = (0xAD70, 2)
    simulate_batched(models[:2], 1000, 250, key, strict=False, adjoint=True)  # warm-up
    torch.cuda.synchronize()
    est = estimate_group_bytes(models, np.full(n_scenes, nf), np.full(n_scenes, nb), key,
                               bytes_per_record=320.0, bytes_per_query=600.0) * 1.3
    groups = plan_groups(est, 0.5 * free_device_bytes())
    tf = tk = ta = 0.0
    for a, b in groups:
        ids = range(a, b)
        # targets at a tenth of the photons, outside the timed region
        rt = simulate_batched(tmodels[a:b], nf // 10, nb // 10, (0x7A26, 7), scene_ids=ids, win_len=1000,
                              strict=False)
        T = [rt.gated(s, tmodels[a + s].gate(1000)).clone() for s in range(b - a)]
        del rt
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = simulate_batched(models[a:b], nf, nb, key, scene_ids=ids, win_len=1000, strict=False)
        torch.cuda.synchronize()
        tf += time.perf_counter() - t0
        del r
        t0 = time.perf_counter()
        res = simulate_batched(models[a:b], nf, nb, key, scene_ids=ids, win_len=1000, strict=False, adjoint=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        d = torch.zeros_like(res.waveforms)
        for s in range(b - a):
            m = models[a + s]
            g0, gl = m.gate(1000)
            o = int(res.offsets[s])
            d[o:o + 4 * m.n_bins].view(4, m.n_bins)[:, g0:g0 + gl] = 2 * (res.waveform(s)[:, g0:g0 + gl] - T[s])
        acc, direct = adjoint.adjoint_sums(res, d)
        derivs = [adjoint.scene_derivatives(m, th[a + s]) for s, m in enumerate(models[a:b])]
        g = adjoint.gradients_from_sums(acc, direct, derivs)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tk += t1 - t0
        ta += t2 - t1
        del res, d
        if not np.all(np.isfinite(g)):
            raise RuntimeError("non-finite adjoint gradient")
        print(f"adjoint scenes {a}-{b} of {n_scenes} done", file=sys.stderr, flush=True)
    return {"workload": f"first {n_scenes} Latin-hypercube sweep scenes (seed 1), {nf} forward + {nb} backward "
                        f"photons each, MSE over the 1000-bin gate against targets from parameters perturbed by "
                        f"U[-10%,10%] (seed 2), in {len(groups)} memory-bounded groups",
            "forward_s": tf, "forward_kept_s": tk, "adjoint_s": ta,
            "ratio": (tk + ta) / tf, "ratio_with_depth_difference": (tk + ta + tf) / tf,
            "depth_difference": "one-sided, common random numbers: one more forward pass",
            "parameters": list(adjoint.PARAMS), "groups": len(groups),
            "forward_paths_per_s": n_scenes * (nf + nb) / tf}


def bench_adjoint_fd(n_scenes: int, replicas: int, nf: int = 100000, eps: float = 5e-2, group: int = 128,
                     first: int = 0, loss: str = "mse", params: str = "abg"):
    """Adjoint a, b, g gradients against central differences with common
    random numbers, per sweep scene, over many independent replicas.

    Scene s (LHS seed 1) with its MSE target (parameters perturbed by
    U[-10 %, 10 %], seed 2; 4 x nf photons in 4 averaged runs) is replicated
    `replicas` times, each replica its own Philox stream, nf forward and nf / 4
    backward photons, in batched runs of `group` replicas.  Each replica's
    adjoint gradient is paired with its own central difference at relative
    step eps (same streams on both sides).  Reported per parameter: the mean
    adjoint gradient, the mean difference, their relative gap, and the
    standard error of the paired difference relative to the mean difference.
    Both estimators are unbiased for the gradient of the expected loss up to
    O(eps^2) curvature and O(1/N) self-interaction terms, so the gap measures
    the adjoint's error within that standard error.

    loss="linear" replaces the MSE by sum 2 (W_ref - T) W, W_ref the scene's
    mean waveform from four separate runs of 4 x nf photons: the MSE
    linearised at the mean.  Its expected value is linear in the expected
    waveform, so its gradient has no variance term.  The MSE's expected
    value also holds the waveform's variance, whose derivative the CRN
    difference sees and the path-replay gradient (samples held fixed) does
    not; where a scene's gated signal is carried by few heavy contributions
    (a floor return inside the gate at 1e5 photons) that term is large."""
    import numpy as np
    import torch

    from pyalbsim_amd import adjoint
    from pyalbsim_amd.execution.batched import simulate_batched
    from pyalbsim_amd.sweep import hypercube

    nb, wl, key = nf // 4, 1000, (5, 6)
    th_all = np.array([[x.params[c] for c in adjoint.PARAMS] for x in hypercube(256, 1)[:first + n_scenes]])
    pert_all = th_all * (1 + np.random.default_rng(2).uniform(-0.1, 0.1, th_all.shape))
    pert_all[:, 2] = np.clip(pert_all[:, 2], 0.5, 0.97)
    cols = tuple((c, n) for c, n in ((0, "a"), (1, "b"), (2, "g")) if n in params)

    def losses(theta, ids, T, adj, R=None):
        models = [adjoint.model_from_theta(theta)] * len(ids)
        res = simulate_batched(models, nf, nb, key, scene_ids=ids, win_len=wl, strict=False, adjoint=adj)
        m = models[0]
        g0, gl = m.gate(wl)
        L = np.zeros(len(ids))
        d = torch.zeros_like(res.waveforms) if adj else None
        for i in range(len(ids)):
            w = res.gated(i, m.gate(wl))
            L[i] = float(((w - T) ** 2).sum()) if R is None else float((R * w).sum())
            if adj:
                o = int(res.offsets[i])
                d[o:o + 4 * m.n_bins].view(4, m.n_bins)[:, g0:g0 + gl] = 2 * (w - T) if R is None else R
        if not adj:
            return L, None
        acc, direct = adjoint.adjoint_sums(res, d)
        derivs = [adjoint.scene_derivatives(m, theta)] * len(ids)
        return L, adjoint.gradients_from_sums(acc, direct, derivs)

    out = {"workload": f"first {n_scenes} Latin-hypercube sweep scenes (seed 1), MSE over the {wl}-bin gate against "
                       f"targets from parameters perturbed by U[-10%,10%] (seed 2); {replicas} independent replicas "
                       f"per scene ({nf} forward + {nb} backward photons each, batched {group} at a time); central "
                       f"differences with common random numbers at relative step {eps}; loss {loss}",
           "scenes": []}
    t0 = time.perf_counter()
    for s in range(first, first + n_scenes):
        th = th_all[s]
        tm = [adjoint.model_from_theta(pert_all[s])] * 4
        rt = simulate_batched(tm, 4 * nf, nf, (77, 78), win_len=wl, strict=False)
        T = sum(rt.gated(r, tm[r].gate(wl)) for r in range(4)) / 4
        del rt
        R = None
        if loss == "linear":
            mm = [adjoint.model_from_theta(th)] * 4
            rr = simulate_batched(mm, 4 * nf, nf, (79, 80), win_len=wl, strict=False)
            R = 2 * (sum(rr.gated(r, mm[r].gate(wl)) for r in range(4)) / 4 - T)
            del rr
        G, FD = [], {c: [] for c, _ in cols}
        g0, gs = 0, group
        while g0 < replicas:
            ids = range(1 + s * replicas + g0, 1 + s * replicas + min(replicas, g0 + gs))
            try:
                gr = losses(th, ids, T, True, R)[1]
                fdr = {}
                for c, _ in cols:
                    tp, tq = th.copy(), th.copy()
                    tp[c] *= 1 + eps
                    tq[c] *= 1 - eps
                    fdr[c] = (losses(tp, ids, T, False, R)[0] - losses(tq, ids, T, False, R)[0]) / (2 * eps * th[c])
            except torch.OutOfMemoryError:
                # a turbid scene's kept adjoint state (neighbour rows) outgrew the
                # device at this group size: the same replicas again, half as many
                # per batched run (their streams, and so their values, are unchanged)
                if gs == 1:
                    raise
                torch.cuda.empty_cache()
                gs //= 2
                continue
            G.append(gr)
            for c, _ in cols:
                FD[c].append(fdr[c])
            g0 += len(ids)
            torch.cuda.synchronize()
            print(f"adjoint vs differences: scene {s}, replicas {g0} of {replicas} ({gs} per run)",
                  file=sys.stderr, flush=True)
        G = np.concatenate(G)
        row = {"scene": s, "theta": [float(v) for v in th]}
        for c, name in cols:
            fd = np.concatenate(FD[c])
            d = G[:, c] - fd
            mf = float(fd.mean())
            row[name] = {"adjoint": float(G[:, c].mean()), "difference": mf,
                         "relative_gap": float(d.mean() / abs(mf)),
                         "relative_se": float(d.std(ddof=1) / np.sqrt(len(d)) / abs(mf)),
                         "adjoint_relative_sd_per_replica": float(G[:, c].std(ddof=1) / abs(G[:, c].mean())),
                         "difference_relative_sd_per_replica": float(fd.std(ddof=1) / abs(mf))}
        out["scenes"].append(row)
        print(json.dumps({"adjoint_vs_differences_scene": row}), file=sys.stderr, flush=True)
    out["seconds"] = time.perf_counter() - t0
    return out


if __name__ == "__main__":
    sys.exit(main())
