#!/usr/bin/env python3
"""Benchmark of the explicit step on one MI355X.

Workload: 125-sphere gravity settling -- spheres of radius 1e-3 m with about
8k jittered-lattice nodes each (1.0M nodes, 92.6M directed bonds), delta = 3h,
walls on six sides, gravity, FP64, dt = 1e-8 s.  Each sphere starts with a
seeded rigid velocity (+-0.5 m/s per axis) plus a seeded per-node velocity
(+-0.01 m/s): without it every node of a sphere shares one displacement and
every bond has exactly zero stretch, which is not a representative load.  One "step" is one full
explicit step of the library (Verlet position half, contact list, bond force
+ break + contact + wall + gravity + Verlet velocity half in one kernel).

value: bond-updates/s over the timed steps (directed bonds of movable
particles per step, broken ones included) with all state resident in HBM.
roofline: the bond-force kernel's algorithmic bytes per launch
(5 B per directed bond + 205 B per node: positions and volume 32, displacement
24, force 24, particle id 4, external force 24, acceleration and velocity
read and written 96, clamp flag 1 -- after its bond sum the kernel adds each
node's contact, wall, external and torque terms and takes the Verlet
velocity half) over its average duration, measured
with HIP events recorded on the library's stream around the bond kernel in
the timed steps themselves (profiled steps launch one by one; the same steps
replayed from graphs without marks are timed beside them and logged).  A
short window after the timed steps marks every phase for the breakdown.
fp64_issue: FP64 lane-instructions per second of the same kernel launch
(wave-instruction count per launch from the committed rocprofv3
SQ_INSTS_VALU_*_F64 summary, times 64 lanes, over the live-measured average
duration) against the FP64 vector issue peak.
cpu_baseline: the CPU oracle (same algorithm, OpenMP threads emulating the
reference's MPI ranks) on a bounded sample of the same workload, all host
threads up to 16 (value) and one thread (single_core), with the host CPU model.

batched: 256 independent two-sphere impacts (node counts per sphere
log-uniform on [500, 20k], speed U(5, 50) m/s, E U(20, 100) GPa with s0 and
contact stiffness per instance) admitted into one job pool and stepped
together, against the same jobs run one after another through the plain
interface.  sims/hour counts whole jobs (context or pool set-up, upload,
steps, and the read-back of every result array: per particle disp, CurrPos,
vel, acc, force and intact bonds, and the wall planes and reaction); the
digests that compare every job's results with its solo run are formed after
the clock stops, on both sides.
serving: a seeded Poisson stream of aggregate-breakage jobs (2-64 spheres and
cubes of 2k-8k nodes each, step budgets drawn from 5k-50k and scaled by
--serve-step-scale) through continuous batching with admission by device
footprint, at several arrival rates; p50/p99 latency and sims/hour next to a
FIFO queue of sequential plain-GPU runs with the same arrivals (service times
measured); a served job is done when its results are read back.
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


def traffic_bytes(nodes, bonds, kernel):
    """HBM bytes per bond-kernel launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary, when it was taken on this same workload
    with the same kernel shape."""
    path = os.path.join(ROOT, "profiles", "k_bond_gravity_settling_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("nodes") != nodes or t.get("movable_bonds") != bonds:
        return None
    # every shape was timed by the autotuner in the profiled run
    return t.get("per_kernel_bytes", {}).get(kernel)


# FP64 vector issue peak: 256 CUs x 4 SIMDs x 16 lanes per clock x 2.4 GHz
# (78.6 TFLOP/s counts each FMA lane-instruction as two flops)
FP64_LANE_INSTR_PEAK = 256 * 64 * 2.4e9


def fp64_instr(kernel):
    """FP64 wave-instructions per launch of `kernel` from the committed
    rocprofv3 SQ_INSTS_VALU_*_F64 summary (same workload and shape)."""
    path = os.path.join(ROOT, "profiles", "k_bond_gravity_settling_fp64.json")
    try:
        with open(path) as f:
            return json.load(f)["per_kernel_fp64_wave_instr"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def shape_name(k, nodes, tol=False):
    """Kernel symbol of bond shape k (pools below 2^27 node slots address
    node records with 32-bit offsets; tol: the tolerance-arithmetic form)."""
    from peridem_amd import abi
    u, pipe, staged = abi.bond_shapes()[k]
    t = ", true" if tol else ", false"
    if staged:
        return f"pd::k_bond_lds<{u}, {pipe}{t}>"
    off32 = "true" if nodes < (1 << 27) else "false"
    return f"pd::k_bond_t<{u}, {pipe}, {off32}{t}>"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def gpu_state_system(s, e, first_step):
    """The workload as the GPU holds it now: every particle's disp, vel, acc
    and intact bonds and the wall planes (what a tc_/wall_ checkpoint holds)."""
    from peridem_amd import gen
    parts = []
    for p, P in enumerate(s.particles):
        st = e.state(p)
        Q = gen.Particle(**P.__dict__)
        Q.disp, Q.vel, Q.acc = st["disp"], st["vel"], st["acc"]
        Q.conn = e.connectivity(p)
        parts.append(Q)
    lrtb, _ = e.wall()
    return gen.System(**{**s.__dict__, "particles": parts, "wall_lrtb": tuple(lrtb), "first_step": first_step})


def cpu_baseline(s, nspheres, steps, threads, steps_1core, mid=None, mid_step=0):
    """The oracle on the first `nspheres` spheres of the workload `s`, on
    `threads` host threads and (a shorter window) on one; with `mid`, the same
    window again from that GPU-produced mid-run state."""
    from oracle import oracle
    from peridem_amd import gen
    if nspheres < len(s.particles):
        s = gen.System(**{**s.__dict__, "particles": s.particles[:nspheres]})
        if mid is not None:
            mid = gen.System(**{**mid.__dict__, "particles": mid.particles[:nspheres]})

    def timed(ranks, n):
        o = oracle.engine(s, ranks=ranks)
        o.step(2)
        t0 = time.perf_counter()
        o.step(n)
        dt = time.perf_counter() - t0
        o.close()
        return s.movable_bonds() * n / dt, dt

    v, dt = timed(threads, steps)
    out = {"value": v, "unit": "bond-updates/s", "cores": threads, "kind": "port",
           "sample": f"{steps} steps (after 2 untimed) of {len(s.particles)} of the 125 spheres "
                     f"({s.nodes} nodes, {s.bonds} directed bonds), the oracle with {threads} "
                     f"OpenMP threads standing in for the reference's MPI ranks, {dt:.1f} s",
           "host": f"{cpu_model()}, {os.cpu_count()} logical CPUs visible"}
    if steps_1core:
        v1, dt1 = timed(1, steps_1core)
        out["single_core"] = {"value": v1, "sample": f"{steps_1core} steps of the same spheres on one "
                                                     f"thread, {dt1:.1f} s"}
    if mid is not None:
        s = mid
        vm, dtm = timed(threads, steps)
        out["mid_run"] = {"value": vm, "sample": f"{steps} steps (after 2 untimed) from the GPU's state after "
                                                 f"{mid_step} steps, {threads} threads, {dtm:.1f} s"}
    return out


def batched(args):
    from peridem_amd import gen, serve
    t0 = time.perf_counter()
    systems = gen.impact_batch(args.batch_jobs)
    nodes = sum(x.nodes for x in systems)
    bonds = sum(x.movable_bonds() for x in systems)
    log(f"[bench] batched: {len(systems)} impacts, {nodes} nodes, {bonds} directed bonds "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    wall_b, dev_ms, dig_b, adm = serve.run_batch(systems, args.batch_steps, digests=True)
    log(f"[bench] batched: {wall_b:.2f} s wall ({adm:.2f} s admission), "
        f"{dev_ms / args.batch_steps:.3f} ms device per step")
    if args.no_sequential:   # profiling the batch alone
        return {"workload": "batched impacts (no sequential comparison)", "jobs": len(systems),
                "steps_per_job": args.batch_steps, "nodes": nodes, "directed_bonds": bonds,
                "batch_wall_s": wall_b, "batch_ms_per_step": dev_ms / args.batch_steps,
                "batch_bond_updates_per_s": bonds * args.batch_steps / (dev_ms * 1e-3)}
    t_seq, dig_s = serve.run_sequential(systems, args.batch_steps, digests=True)
    wall_s = sum(t_seq)
    same = sum(a == b for a, b in zip(dig_b, dig_s))
    log(f"[bench] sequential: {wall_s:.2f} s wall; {same}/{len(systems)} results identical")
    return {"workload": f"{len(systems)} two-sphere impacts, nodes per sphere log-uniform on "
                        f"[500, 20k], v U(5,50) m/s, E U(20,100) GPa, {args.batch_steps} steps each, "
                        f"FP64, dt=1e-8",
            "jobs": len(systems), "steps_per_job": args.batch_steps, "nodes": nodes,
            "directed_bonds": bonds,
            "sims_per_hour": len(systems) / wall_b * 3600.0,
            "sequential_sims_per_hour": len(systems) / wall_s * 3600.0,
            "speedup_vs_sequential": wall_s / wall_b,
            "batch_wall_s": wall_b, "batch_admission_s": adm, "sequential_wall_s": wall_s,
            "batch_ms_per_step": dev_ms / args.batch_steps,
            "batch_bond_updates_per_s": bonds * args.batch_steps / (dev_ms * 1e-3),
            "results_identical_to_solo": f"{same}/{len(systems)}"}


def config0_latency(args):
    """The two-sphere impact of the reference's 3-D example at its own size
    (R = 1e-3 m, h = 1e-4 m: 4,169 nodes per sphere, A at 20 m/s toward a
    fixed B, contact from about step 50), all of its steps in one call: the
    device time per step of a pool far too small to fill the GPU, where a step
    is bound by the latency of its kernels."""
    from peridem_amd import abi, gen
    s = gen.two_sphere_impact()
    e = abi.gpu_engine(s)
    try:
        t0 = time.perf_counter()
        e.step(args.config0_steps)
        wall = time.perf_counter() - t0
        dev = e.last_step_ms()
        c = e.counters()
    finally:
        e.close()
    out = {"workload": "two spheres R=1e-3 m, h=1e-4 m (4169 nodes each), A at 20 m/s toward fixed B, dt=1e-8",
           "nodes": s.nodes, "directed_bonds": s.bonds, "steps": args.config0_steps,
           "device_ms_per_step": dev / args.config0_steps, "wall_s": wall,
           "pair_computations": c["pair_computations"], "alive_bonds": c["alive_bonds"]}
    log(f"[bench] config 0: {args.config0_steps} steps in {wall:.3f} s wall, "
        f"{out['device_ms_per_step'] * 1e3:.1f} us device per step")
    return out


# wave slots of the bond kernel on MI355X: 256 CUs x 4 SIMDs x 2 waves
WAVE_SLOTS = 256 * 4 * 2


def serving(args):
    from peridem_amd import abi, gen, serve
    jobs = []
    for i in range(args.serve_jobs):
        sysm = gen.aggregate_job(i)
        jobs.append((sysm, max(1, int(round(sysm.meta["steps"] * args.serve_step_scale)))))
    nodes = [x.nodes for x, _ in jobs]
    log(f"[bench] serving: {len(jobs)} jobs, {min(nodes)}-{max(nodes)} nodes, steps "
        f"{min(n for _, n in jobs)}-{max(n for _, n in jobs)}")
    dev_seq = []
    if args.no_sequential:   # profiling the served pool alone
        t_seq, dig_s = [float("nan")] * len(jobs), [None] * len(jobs)
    else:
        t_seq, dig_s = serve.run_sequential([x for x, _ in jobs], [n for _, n in jobs], digests=True,
                                            device_ms=dev_seq)
        log(f"[bench] serving: sequential service {sum(t_seq):.2f} s total, {sum(dev_seq) * 1e-3:.2f} s of it "
            f"stepping on the device")
    # how much of the GPU one job fills: its bond kernel has one wavefront per
    # 64 node slots and the GPU holds WAVE_SLOTS of them at two per SIMD
    fill = [x.nodes / 64.0 / WAVE_SLOTS for x, _ in jobs]
    # pool sized by a device-memory budget; a job is admitted when its footprint fits
    fps = [abi.footprint(x) for x, _ in jobs]
    budget = args.serve_pool_gib * (1 << 30)
    tot_b = sum(f[2] for f in fps)
    frac = min(1.0, budget / tot_b)
    cap = dict(max_nodes=max(max(f[0] for f in fps), int(sum(f[0] for f in fps) * frac)),
               max_bond_entries=max(max(f[1] for f in fps), int(sum(f[1] for f in fps) * frac)),
               max_particles=sum(len(x.particles) for x, _ in jobs) + 1, max_jobs=len(jobs))
    out = {"workload": f"Poisson stream of {len(jobs)} aggregate-breakage jobs (2-64 seeded spheres and "
                       f"cubes of 2k-8k nodes, walls, self-contact), step budgets 5k-50k scaled by "
                       f"{args.serve_step_scale}, FP64, dt=1e-8",
           "jobs": len(jobs), "pool_budget_gib": args.serve_pool_gib,
           "pool_node_slots": cap["max_nodes"], "pool_bond_entries": cap["max_bond_entries"],
           "sequential_sims_per_hour": len(jobs) / sum(t_seq) * 3600.0,
           "sequential_device_s": sum(dev_seq) * 1e-3, "sequential_wall_s": sum(t_seq),
           "job_nodes": [x.nodes for x, _ in jobs], "job_steps": [n for _, n in jobs],
           "job_gpu_fill": {"definition": f"nodes / 64 / {WAVE_SLOTS} (the bond kernel's wavefronts over the "
                                          f"wave slots of 256 CUs x 4 SIMDs x 2)",
                            "min": min(fill), "median": float(sorted(fill)[len(fill) // 2]), "max": max(fill),
                            "jobs_at_least_one_gpu": sum(f >= 1.0 for f in fill)},
           "rates": []}
    for rate in args.serve_rates:
        arr = serve.poisson_arrivals(len(jobs), rate, seed=int(rate * 1000) + 1)
        srv = serve.Server(chunk=args.serve_chunk, **cap)
        t0 = time.perf_counter()
        try:
            done = srv.serve([(a, x, n) for a, (x, n) in zip(arr, jobs)])
        finally:
            srv.close()
        span = max(d.finished for d in done) - float(arr[0])
        lat = [d.latency for d in sorted(done, key=lambda d: d.rid)]
        same = sum(d.digest == dig_s[d.rid] for d in done)
        fifo = serve.fifo_latency(arr, t_seq)
        t = 0.0
        for a, sv in zip(arr, t_seq):
            t = max(t, a) + sv
        seq_span = t - float(arr[0])
        rec = {"rate_jobs_per_s": rate, "served": serve.latency_stats(lat),
               "sequential_fifo": serve.latency_stats(fifo),
               "sims_per_hour": len(jobs) / span * 3600.0,
               "sequential_sims_per_hour_same_arrivals": len(jobs) / seq_span * 3600.0,
               "peak_resident_jobs": srv.peak_live, "results_identical_to_solo": f"{same}/{len(jobs)}",
               "pool_steps": srv.steps_run, "device_s": srv.device_ms * 1e-3,
               "device_gain": sum(dev_seq) / max(srv.device_ms, 1e-9),
               "wall_s": time.perf_counter() - t0}
        log(f"[bench] serving rate {rate}/s: p50 {rec['served']['p50_s']:.3f} s, p99 "
            f"{rec['served']['p99_s']:.3f} s (sequential FIFO p50 {rec['sequential_fifo']['p50_s']:.3f}, "
            f"p99 {rec['sequential_fifo']['p99_s']:.3f}); {rec['sims_per_hour']:.0f} sims/h; "
            f"peak {srv.peak_live} resident; {same}/{len(jobs)} identical; {srv.steps_run} pool steps, "
            f"{srv.device_ms * 1e-3:.2f} s on the device (alone: {sum(dev_seq) * 1e-3:.2f} s)")
        out["rates"].append(rec)
    if args.serve_phases:
        # where the served pool's device time goes: the highest rate once
        # more with a mark around every phase of every step (steps launched
        # one by one, so this run is not one of the timed ones above)
        rate = max(args.serve_rates)
        arr = serve.poisson_arrivals(len(jobs), rate, seed=int(rate * 1000) + 1)
        srv = serve.Server(chunk=args.serve_chunk, digests=False, **cap)
        srv.pool.set_profiling(abi.PD_PROFILE_PHASES)
        try:
            srv.serve([(a, x, n) for a, (x, n) in zip(arr, jobs)])
            ph = srv.pool.profile()
        finally:
            srv.close()
        tot = sum(v["ms"] for v in ph.values())
        out["phases_at_rate"] = {"rate_jobs_per_s": rate, "pool_steps": srv.steps_run,
                                 "device_ms": {k: v["ms"] for k, v in ph.items()},
                                 "share": {k: v["ms"] / tot for k, v in ph.items()}}
        log(f"[bench] serving phases at {rate}/s over {srv.steps_run} pool steps: " +
            ", ".join(f"{k} {v['ms'] * 1e-3:.2f} s ({v['ms'] / tot:.1%})" for k, v in ph.items()))
    return out


def crush(args):
    """The crushing aggregate at full size: 512 mixed cubes, spheres and convex
    polyhedra (2k-12k nodes each, h = 1e-4 m, self-contact, friction,
    damping, wall reaction), the top wall moving down at 5 m/s until 30% strain
    of the box height, then all wall speeds set to zero
    (set_zero_wall_speed_timestep, timeloop.h:447-462) and the packing left to
    relax.  Every chunk of steps records the top wall's travel, the reaction
    on the top and bottom walls (contact.h:255-272: -sum V f over the nodes in
    contact; the reaction of one step, averaged over samples taken every
    `crush_sample` steps of the chunk), the fraction of directed bonds broken
    and the particle pairs that passed the box test per step."""
    from peridem_amd import abi, gen
    s = gen.crushing_aggregate(512)
    H0 = s.wall_lrtb[5] - s.wall_lrtb[2]
    speed = -s.wall_speed[5]
    n30 = int(round(0.3 * H0 / (speed * s.dt)))
    end = n30 + args.crush_relax
    log(f"[crush] {len(s.particles)} particles, {s.nodes} nodes, {s.bonds} directed bonds; box height "
        f"{H0:.4e} m; 30% strain after {n30} steps, then {args.crush_relax} steps at rest")
    e = abi.gpu_engine(s)
    rows, t, dev_ms, phase_rows = [], 0, 0.0, []
    c0 = e.counters()
    t_start = time.perf_counter()
    try:
        while t < end:
            if t == n30:
                e.set_wall_speed([0.0] * 6)
            n = min(args.crush_chunk, (n30 if t < n30 else end) - t)
            # the reaction of single steps, sampled every `sample` steps of the chunk
            top, bottom, done = [], [], 0
            while done < n:
                m = min(args.crush_sample, n - done)
                # every crush_phase_every steps, the first sample of the
                # chunk carries a mark around every phase of each step
                marked = args.crush_phase_every > 0 and done == 0 and t > 0 and t % args.crush_phase_every == 0
                if marked:
                    e.set_profiling(abi.PD_PROFILE_PHASES)
                e.step(m)
                dev_ms += e.last_step_ms()
                if marked:
                    ph = e.profile()
                    e.set_profiling(0)
                    phase_rows.append({"step": t, "steps": m, "pairs_per_step": None,
                                       "ms": {k: v["ms"] / max(v["launches"], 1) for k, v in ph.items()}})
                done += m
                _, reac = e.wall()
                top.append(float(reac[5][2]))
                bottom.append(float(reac[2][2]))
            t += n
            lrtb, _ = e.wall()
            c = e.counters()
            row = {"step": t, "wall_travel_m": s.wall_lrtb[5] - lrtb[5],
                   "strain": (s.wall_lrtb[5] - lrtb[5]) / H0,
                   "reaction_top_z_N": sum(top) / len(top), "reaction_top_z_max_N": max(top, key=abs),
                   "reaction_bottom_z_N": sum(bottom) / len(bottom),
                   "broken_fraction": 1.0 - c["alive_bonds"] / c["total_bonds"],
                   "pairs_per_step": (c["pair_computations"] - c0["pair_computations"]) / n,
                   "ms_per_step": dev_ms / max(t, 1), "list_capacity": c["list_capacity"]}
            if phase_rows and phase_rows[-1]["pairs_per_step"] is None:
                phase_rows[-1]["pairs_per_step"] = row["pairs_per_step"]
            c0 = c
            rows.append(row)
            log(f"[crush] step {t}: strain {row['strain']:.4f}, top reaction mean {row['reaction_top_z_N']:.4e} N "
                f"(largest {row['reaction_top_z_max_N']:.4e}), bottom {row['reaction_bottom_z_N']:.4e} N, broken "
                f"{row['broken_fraction']:.3e}, pairs/step {row['pairs_per_step']:.0f}, list capacity "
                f"{row['list_capacity']}")
    finally:
        cnt = e.counters()
        e.close()
    wall_s = time.perf_counter() - t_start
    out = {"mode": "crush", "workload": "crushing aggregate, 512 cubes/spheres/convex polyhedra of 2k-12k "
                                        "nodes (uniform) placed at random (bounding spheres 35% of the box), h=1e-4 m, "
                                        "delta=3h, self-contact, contact stiffness of 2e9 Pa, top wall -5 m/s to "
                                        "30% strain then at rest, FP64, dt=1e-8",
           "nodes": s.nodes, "directed_bonds": s.bonds, "box_height_m": H0, "steps_to_30pct": n30,
           "relax_steps": args.crush_relax, "steps": t, "wall_s": wall_s,
           "device_s": dev_ms * 1e-3, "bond_updates_per_s": s.movable_bonds() * t / (dev_ms * 1e-3),
           "list_regrows": cnt["list_regrows"], "cell_rebuilds": cnt["cell_rebuilds"], "trace": rows,
           "phases": phase_rows}
    print(json.dumps(out), flush=True)
    if args.crush_md:
        with open(args.crush_md, "w") as f:
            f.write("| step | top wall travel (m) | strain | top wall reaction z, mean (N) | largest (N) "
                    "| bottom wall reaction z, mean (N) | broken bonds | pairs/step |\n"
                    "|---|---|---|---|---|---|---|---|\n")
            for r in rows:
                f.write(f"| {r['step']} | {r['wall_travel_m']:.4e} | {r['strain']:.4f} | "
                        f"{r['reaction_top_z_N']:.4e} | {r['reaction_top_z_max_N']:.4e} | "
                        f"{r['reaction_bottom_z_N']:.4e} | {r['broken_fraction']:.4e} | "
                        f"{r['pairs_per_step']:.0f} |\n")
            if phase_rows:
                names = list(phase_rows[0]["ms"])
                f.write("\n| step | pairs/step | " + " | ".join(f"{k} (ms)" for k in names) + " | step (ms) |\n"
                        "|---|---|" + "---|" * len(names) + "---|\n")
                for r in phase_rows:
                    f.write(f"| {r['step']} | {r['pairs_per_step']:.0f} | " +
                            " | ".join(f"{r['ms'][k]:.4f}" for k in names) +
                            f" | {sum(r['ms'].values()):.4f} |\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle-s", type=float, default=0.5,
                    help="untimed steps for this long after the first (tuning) step, before the warm-up")
    ap.add_argument("--phase-steps", type=int, default=40,
                    help="steps after the timed ones with a mark around every phase")
    ap.add_argument("--spheres", type=int, default=125)
    ap.add_argument("--cpu-spheres", type=int, default=125)
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-steps-1core", type=int, default=8,
                    help="steps of the single-thread CPU sample (0: skip it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--config0-steps", type=int, default=2000,
                    help="steps of the two-sphere example timed for its per-step latency (0: skip)")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched and served legs without the one-by-one comparison runs (for profiling)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tolerance", action="store_true",
                    help="skip the tolerance-arithmetic line")
    ap.add_argument("--serve-jobs", type=int, default=24)
    ap.add_argument("--serve-rates", type=lambda v: [float(x) for x in v.split(",")],
                    default=[1.0, 2.0, 5.0, 10.0, 20.0, 50.0])
    ap.add_argument("--serve-step-scale", type=float, default=0.05)
    ap.add_argument("--serve-pool-gib", type=float, default=1.0)
    ap.add_argument("--serve-chunk", type=int, default=256)
    ap.add_argument("--serve-phases", action="store_true",
                    help="serve the highest rate once more with every phase of every step timed")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (a plain copy "
                         "kernel of known size that calibrates FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--bond-shape", type=int, default=-1,
                    help="force this bond-kernel shape instead of the autotuner's choice "
                         "(profiling passes pin one so that every pass times the same kernel)")
    ap.add_argument("--no-headline", action="store_true",
                    help="run only the batched and served legs (their JSON, not the benchmark line)")
    ap.add_argument("--crush", action="store_true",
                    help="instead of the benchmark: the crushing aggregate at full size to 30%% "
                         "strain and a relaxation, printing its force-displacement trace")
    ap.add_argument("--crush-chunk", type=int, default=2000)
    ap.add_argument("--crush-sample", type=int, default=100,
                    help="steps between the wall reactions averaged over a chunk")
    ap.add_argument("--crush-relax", type=int, default=10000)
    ap.add_argument("--crush-md", default="", help="also write the trace as a markdown table here")
    ap.add_argument("--crush-phase-every", type=int, default=32000,
                    help="every this many steps, one sample of steps with a mark around every phase (0: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library steps on one GPU; sharding over several is a design "
                                   "note only (particles shard by spatial block with a surface "
                                   "halo exchange)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.crush:
        crush(args)
        return
    if args.no_headline:   # the batched and served legs alone (profiling, long serving sweeps)
        out = {"mode": "batched/served legs only",
               "batched": None if args.no_batch else batched(args),
               "serving": None if args.no_serve else serving(args)}
        print(json.dumps(out), flush=True)
        return

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
    if args.bond_shape >= 0:
        e.set_bond_shape(args.bond_shape)
    log(f"[bench] committed to the GPU in {time.perf_counter() - t0:.1f} s")
    # the first step tunes the bond kernel and builds the candidate lists;
    # then untimed steps for settle_s seconds bring the GPU to the clocks it
    # runs a long simulation at, before the warm-up steps
    e.step(1)
    settled, t0 = 1, time.perf_counter()
    while time.perf_counter() - t0 < args.settle_s:
        e.step(50)
        settled += 50
    e.step(args.warmup)
    # same number of steps replayed from the step graphs, no marks: logged
    # beside the timed rate so the cost of the marks stays visible
    t0 = time.perf_counter()
    e.step(args.steps)
    graph_s = time.perf_counter() - t0
    # the timed steps carry HIP event marks around the bond kernel, recorded
    # on the library's stream between the kernels they time
    e.set_profiling(abi.PD_PROFILE_BOND)
    c0 = e.counters()
    t0 = time.perf_counter()
    e.step(args.steps)                         # returns after the stream is drained
    elapsed = time.perf_counter() - t0
    c1 = e.counters()
    prof = e.profile()
    dev_ms = e.last_step_ms()
    # the breakdown of a step over its phases, from a short window after the
    # timed one (a mark around every phase)
    e.set_profiling(abi.PD_PROFILE_PHASES)
    e.step(args.phase_steps)
    phases = e.profile()
    # full rebuilds of every node's candidate list, which the timed window
    # does not hold (lists are rebuilt once some node has moved half the
    # skin): a particle flag set to the value it has marks the job for a
    # rebuild in its next step, timed by that step's phase marks
    rebuilds_seen, steps_seen = e.counters()["cell_rebuilds"], e.counters()["steps"]
    rebuild_ms = []
    for _ in range(3):
        e.set_profiling(abi.PD_PROFILE_PHASES)
        e.set_particle_flag(0, abi.PD_PFLAG_MOVABLE, 1)
        e.step(1)
        rebuild_ms.append(e.profile()["contact_list"]["ms"])
    e.set_profiling(0)
    bu = c1["bond_updates"] - c0["bond_updates"]
    log(f"[bench] {args.steps} steps: {elapsed * 1e3:.2f} ms wall, {dev_ms:.2f} ms device "
        f"(unmarked graph replay of as many steps: {graph_s * 1e3:.2f} ms wall)")
    force_ms = prof["bond"]["ms"] / max(prof["bond"]["launches"], 1)
    force_bytes = 5 * Bm + 205 * N
    achieved = force_bytes / (force_ms * 1e-3) / 1e9
    step_ms = elapsed * 1e3 / args.steps
    step_bytes = 5 * Bm + 252 * N
    log(f"[bench] bond kernel over the timed steps: {force_ms:.4f} ms per launch; phases per step over "
        f"{args.phase_steps} more steps (ms): " +
        ", ".join(f"{k}={v['ms'] / max(v['launches'], 1):.4f}" for k, v in phases.items()))
    log(f"[bench] full candidate-list rebuild: {', '.join(f'{x:.3f}' for x in rebuild_ms)} ms "
        f"(contact-list phase of a step that rebuilds every list); {rebuilds_seen} rebuilds in the "
        f"first {steps_seen} steps")
    cnt = e.counters()
    kname = shape_name(e.bond_shape(), N)
    fi = fp64_instr(kname)
    # a GPU-produced mid-run state for the CPU baseline's second window
    mid = None if args.no_cpu else gpu_state_system(s, e, cnt["steps"] + 1)
    mid_step = cnt["steps"]
    tol = None
    if not args.no_tolerance:
        # the same pool in tolerance arithmetic (its shapes tuned again at the
        # first step), timed the same way
        e.set_arithmetic(abi.PD_ARITH_TOLERANCE)
        e.step(args.warmup)
        e.set_profiling(abi.PD_PROFILE_BOND)
        c0 = e.counters()
        t0 = time.perf_counter()
        e.step(args.steps)
        el = time.perf_counter() - t0
        c1 = e.counters()
        pt = e.profile()
        e.set_profiling(0)
        tms = pt["bond"]["ms"] / max(pt["bond"]["launches"], 1)
        ta = force_bytes / (tms * 1e-3) / 1e9
        tk = shape_name(e.bond_shape(), N, tol=True)
        tfi = fp64_instr(tk)
        tol = {"arithmetic": "PD_ARITH_TOLERANCE: |e| and |xi| from v_rsq_f64 with one Newton-Raphson "
                             "step, divisions as products with those reciprocals",
               "tolerance": "per step and node |f - f_ref| <= 1e-12 * c0 * sum_j V_j (a stretch error of "
                            "1e-12); broken set exact for |s - s0| > 1e-12 (tests/test_gpu_tolerance.py)",
               "value": (c1["bond_updates"] - c0["bond_updates"]) / el, "unit": "bond-updates/s",
               "ms_per_step": el * 1e3 / args.steps,
               "step_hbm_fraction": step_bytes / (el / args.steps) / (HBM_PEAK_GBS * 1e9),
               "roofline": {"bound": "hbm", "achieved": ta, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": ta / HBM_PEAK_GBS, "kernel": tk, "bytes_per_launch": force_bytes,
                            "avg_launch_ms": tms, "target_frac": 0.6},
               "fp64_issue": None if tfi is None else {
                   "wave_instr_per_launch": tfi, "frac": tfi * 64 / (tms * 1e-3) / FP64_LANE_INSTR_PEAK}}
        log(f"[bench] tolerance arithmetic: {tol['ms_per_step']:.4f} ms/step, bond kernel {tms:.4f} ms "
            f"({tk}), {ta:.0f} GB/s = {ta / HBM_PEAK_GBS:.3f} of HBM peak")
    e.close()

    if probe is not None:
        import torch
        b = torch.empty_like(probe)
        b.copy_(probe)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {args.copy_probe_mib} MiB device-to-device")

    lat0 = None if args.config0_steps <= 0 else config0_latency(args)
    bat = None if args.no_batch else batched(args)
    srv = None if args.no_serve else serving(args)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: {args.cpu_spheres} spheres x {args.cpu_steps} steps, "
            f"{args.cpu_threads} threads")
        cpu = cpu_baseline(s, args.cpu_spheres, args.cpu_steps, args.cpu_threads, args.cpu_steps_1core,
                           mid=mid, mid_step=mid_step)

    out = {
        "metric": "bond-updates/s & % HBM peak; batched sims/hr + p99 latency vs plain GPU",
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
                   "phase_ms": {k: v["ms"] / max(v["launches"], 1) for k, v in phases.items()},
                   "list_rebuild": {"full_rebuild_ms": sum(rebuild_ms) / len(rebuild_ms),
                                    "samples_ms": rebuild_ms,
                                    "rebuilds_before_forced": rebuilds_seen, "steps_before_forced": steps_seen,
                                    "settle_steps": settled}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_bytes(N, Bm, kname),
                     "kernel": kname, "bytes_per_launch": force_bytes,
                     "avg_launch_ms": force_ms},
        "fp64_issue": None if fi is None else {
            "wave_instr_per_launch": fi, "achieved": fi * 64 / (force_ms * 1e-3) / 1e12,
            "peak": FP64_LANE_INSTR_PEAK / 1e12, "unit": "T FP64 lane-instructions/s",
            "frac": fi * 64 / (force_ms * 1e-3) / FP64_LANE_INSTR_PEAK},
        "tolerance_mode": tol,
        "cpu_baseline": cpu,
        "config0_latency": lat0,
        "batched": bat,
        "serving": srv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
