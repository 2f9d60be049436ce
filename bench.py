#!/usr/bin/env python3
"""Benchmark of the explicit step on one MI355X.

Workload: 125-sphere gravity settling -- spheres of radius 1e-3 m with about
8k jittered-lattice nodes each (1.0M nodes, 92.6M directed bonds), delta = 3h,
walls on six sides, gravity, FP64, dt = 1e-8 s.  Each sphere starts with a
seeded rigid velocity (+-0.5 m/s per axis) plus a seeded per-node velocity
(+-0.01 m/s): without it every node of a sphere shares one displacement and
every bond has exactly zero stretch, which is not a representative load.  One "step" is one full
explicit step of the library (Verlet position half, contact list, bond force
+ break + contact + wall + gravity, Verlet velocity half).

value: bond-updates/s over the timed steps (directed bonds of movable
particles per step, broken ones included) with all state resident in HBM.
roofline: the bond-force kernel's algorithmic bytes per launch
(5 B per directed bond + 84 B per node) over its average duration, measured
with HIP events on the library's stream in a separate profiled run.
cpu_baseline: the CPU oracle (same algorithm, OpenMP threads emulating the
reference's MPI ranks) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


WORKLOAD = dict(sphere_speed=0.5, node_speed=0.01)


def traffic_bytes(nodes, bonds):
    """HBM bytes per k_bond launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary, when it was taken on this same workload."""
    path = os.path.join(ROOT, "profiles", "k_bond_gravity_settling_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("nodes") != nodes or t.get("movable_bonds") != bonds:
        return None
    return t.get("hbm_bytes_per_launch")


def cpu_baseline(nspheres, steps, threads):
    """The oracle on the first `nspheres` spheres of the workload."""
    from oracle import oracle
    from peridem_amd import gen
    s = gen.gravity_settling(nspheres=nspheres, **WORKLOAD)
    o = oracle.engine(s, ranks=threads)
    o.step(2)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    bu = s.movable_bonds() * steps
    o.close()
    return {"value": bu / dt, "unit": "bond-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of {nspheres} of the 125 spheres ({s.nodes} nodes, "
                      f"{s.bonds} directed bonds), oracle with {threads} OpenMP ranks, "
                      f"{dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--spheres", type=int, default=125)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-spheres", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (a plain copy "
                         "kernel of known size that calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "note only (particles shard by spatial block with a surface "
                                   "halo exchange)", "n_gpus": args.gpus}))
        sys.exit(2)

    probe = None
    if args.copy_probe_mib:
        # torch's HIP state is set up before the library's own
        import torch
        probe = torch.ones(args.copy_probe_mib * 1024 * 1024 // 8, dtype=torch.float64,
                           device="cuda")
        torch.cuda.synchronize()
    from peridem_amd import abi, gen
    t0 = time.perf_counter()
    s = gen.gravity_settling(nspheres=args.spheres, **WORKLOAD)
    N, B, Bm = s.nodes, s.bonds, s.movable_bonds()
    log(f"[bench] workload: {args.spheres} spheres, {N} nodes, {B} directed bonds "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    e = abi.gpu_engine(s)
    log(f"[bench] committed to the GPU in {time.perf_counter() - t0:.1f} s")
    e.step(args.warmup)
    c0 = e.counters()
    t0 = time.perf_counter()
    e.step(args.steps)                         # returns after the stream is drained
    elapsed = time.perf_counter() - t0
    c1 = e.counters()
    bu = c1["bond_updates"] - c0["bond_updates"]
    dev_ms = e.last_step_ms()
    log(f"[bench] {args.steps} steps: {elapsed * 1e3:.2f} ms wall, {dev_ms:.2f} ms device")

    e.set_profiling(True)
    e.step(args.profile_steps)
    prof = e.profile()
    e.set_profiling(False)
    force_ms = prof["bond"]["ms"] / max(prof["bond"]["launches"], 1)
    force_bytes = 5 * Bm + 84 * N
    achieved = force_bytes / (force_ms * 1e-3) / 1e9
    step_ms = elapsed * 1e3 / args.steps
    step_bytes = 5 * Bm + 252 * N
    log(f"[bench] phases per step (ms): " +
        ", ".join(f"{k}={v['ms'] / max(v['launches'], 1):.4f}" for k, v in prof.items()))
    cnt = e.counters()
    e.close()

    if probe is not None:
        import torch
        b = torch.empty_like(probe)
        b.copy_(probe)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {args.copy_probe_mib} MiB device-to-device")

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: {args.cpu_spheres} spheres x {args.cpu_steps} steps, "
            f"{args.cpu_threads} threads")
        cpu = cpu_baseline(args.cpu_spheres, args.cpu_steps, args.cpu_threads)

    out = {
        "metric": "bond-updates/s",
        "value": bu / elapsed,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gravity settling, {args.spheres} spheres R=1e-3 m, ~8k "
                               f"jittered-lattice nodes each, delta=3h, 6 walls, gravity, seeded velocities "
                               f"(0.5 m/s rigid + 0.01 m/s per node), dt=1e-8",
                   "nodes": N, "directed_bonds": B, "movable_bonds": Bm,
                   "alive_bonds": cnt["alive_bonds"], "parallelism": "single GPU",
                   "step_hbm_fraction": step_bytes / (step_ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
                   "phase_ms": {k: v["ms"] / max(v["launches"], 1) for k, v in prof.items()}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_bytes(N, Bm),
                     "kernel": "k_bond", "bytes_per_launch": force_bytes,
                     "avg_launch_ms": force_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
