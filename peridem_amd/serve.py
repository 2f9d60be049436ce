"""Batched and served execution of many independent simulations on one GPU.

The reference runs one system per process (run/simulate3d.cpp:142, one
run_timeloop call).  Here a device-resident job pool (pd_reserve /
pd_add_job, include/peridem_gpu.h) holds many systems at once: every step
launch advances all of them, each with its own clock and step budget.

* `run_batch`   admits a fixed set of jobs and steps them to completion.
* `Server`      continuous batching of a request stream: requests are admitted
                in arrival order as job slots and pool memory (the job's
                footprint in node slots and bond entries) free up, stepped in
                chunks, harvested and removed when their budget is used, so
                requests that arrive while others run overlap on the device.
* `run_sequential` / `fifo_latency`  the plain-GPU reference point: each job
                alone through the plain interface, one after another.

A job's results are bitwise those of the same system run alone (the pool
tests check this), so batching changes throughput and latency, not answers.
"""
from __future__ import annotations

import hashlib
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import abi


# host threads laying out jobs (each call is single-threaded for small jobs)
PREPARE_WORKERS = 16


def read_results(read_state, read_conn, read_wall, nparticles):
    """Every result array of a job as read back from the device: per particle
    disp, CurrPos, vel, acc, force and its intact bonds, then the wall planes
    and reaction."""
    out = []
    for p in range(nparticles):
        st = read_state(p)
        out.extend(st[k] for k in ("disp", "currpos", "vel", "acc", "force"))
        out.append(read_conn(p))
    out.extend(read_wall())
    return out


def digest_arrays(arrays):
    """SHA-1 over result arrays: equal digests = bitwise-equal results."""
    h = hashlib.sha1()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def digest_state(read_state, read_conn, read_wall, nparticles):
    return digest_arrays(read_results(read_state, read_conn, read_wall, nparticles))


def pool_capacity(systems, slack=1.0):
    """Pool sizes that hold all `systems` at once (times slack)."""
    fp = [abi.footprint(s) for s in systems]
    return dict(max_nodes=int(sum(f[0] for f in fp) * slack),
                max_bond_entries=int(sum(f[1] for f in fp) * slack),
                max_particles=int(sum(len(s.particles) for s in systems) * slack) + 1,
                max_jobs=max(1, int(len(systems) * slack)))


@dataclass
class Done:
    rid: int
    arrival: float
    admitted: float
    finished: float
    nsteps: int
    nodes: int
    digest: str | None = None
    error: str | None = None      # why the library stopped the job (no results then)

    @property
    def latency(self):
        return self.finished - self.arrival


def run_batch(systems, nsteps, digests=False):
    """Admit all systems as jobs of one pool and step them to completion.

    Returns (wall seconds incl. admission and the read-back of every result
    array, device ms of the stepping, digests, admission seconds).  The
    digests are computed after the clock stops: they verify, they are not part
    of a job."""
    cap = pool_capacity(systems)
    t0 = time.perf_counter()
    pool = abi.Pool(cap["max_nodes"], cap["max_bond_entries"], cap["max_particles"], cap["max_jobs"])
    try:
        budgets = nsteps if isinstance(nsteps, (list, tuple)) else [nsteps] * len(systems)
        # the host layout of every job (pd_job_prepare, thread-safe) on worker
        # threads, placed in arrival order as each layout completes (placing
        # overlaps the later layouts)
        jobs = []
        with ThreadPoolExecutor(max_workers=PREPARE_WORKERS) as ex:
            for b, s, n in zip(ex.map(abi.prepare_job, systems), systems, budgets):
                j = pool.add_prepared(b, s, n)
                b.free()
                if j is None:
                    raise RuntimeError("job does not fit the pool sized for it")
                jobs.append(j)
        t1 = time.perf_counter()
        dev_ms = 0.0
        left = max(budgets)
        while left:
            n = min(left, 1 << 20)
            pool.step(n)
            dev_ms += pool.last_step_ms()
            left -= n
        res = []
        if digests:
            for j, s in zip(jobs, systems):
                res.append(read_results(lambda p, j=j: pool.state(j, p),
                                        lambda p, j=j: pool.connectivity(j, p),
                                        lambda j=j: pool.wall(j), len(s.particles)))
        t2 = time.perf_counter()
        out = [digest_arrays(r) for r in res]
        return t2 - t0, dev_ms, out, t1 - t0
    finally:
        pool.close()


def run_sequential(systems, nsteps, digests=False, device_ms=None):
    """Each system alone through the plain interface, one after another.

    Returns per-job wall seconds (context creation, upload, stepping and the
    read-back of every result array) and digests (computed after each job's
    clock stops).  device_ms, when a list, receives each job's device time of
    its steps (HIP events around the pd_step call)."""
    times, out = [], []
    budgets = nsteps if isinstance(nsteps, (list, tuple)) else [nsteps] * len(systems)
    for s, n in zip(systems, budgets):
        t0 = time.perf_counter()
        e = abi.gpu_engine(s)
        try:
            e.step(n)
            if device_ms is not None:
                device_ms.append(e.last_step_ms())
            res = read_results(e.state, e.connectivity, e.wall, len(s.particles)) if digests else None
        finally:
            e.close()
        times.append(time.perf_counter() - t0)
        out.append(digest_arrays(res) if digests else None)
    return times, out


def fifo_latency(arrivals, service):
    """Job latencies of a single FIFO server with the given service times."""
    t, lat = 0.0, []
    for a, s in zip(arrivals, service):
        t = max(t, a) + s
        lat.append(t - a)
    return lat


def poisson_arrivals(n, rate, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


@dataclass
class Server:
    """Continuous batching over one job pool.

    chunk: most steps per pd_step call; a call never runs past the earliest
    budget end among the resident jobs, so finished jobs leave (and waiting
    ones enter) at the next call boundary.
    prepare_ahead: requests laid out on the host (pd_job_prepare, on a worker
    thread) while the pool steps; admission then only places them."""
    max_nodes: int
    max_bond_entries: int
    max_particles: int
    max_jobs: int
    chunk: int = 256
    digests: bool = True
    prepare_ahead: int = 2
    pool: abi.Pool = field(init=False)
    queue: deque = field(init=False, default_factory=deque)
    live: dict = field(init=False, default_factory=dict)
    done: list = field(init=False, default_factory=list)
    results: list = field(init=False, default_factory=list)   # read-back arrays per done job, until serving ends

    def __post_init__(self):
        self.pool = abi.Pool(self.max_nodes, self.max_bond_entries, self.max_particles, self.max_jobs)
        self.steps_run = 0
        self.device_ms = 0.0
        self.peak_live = 0
        self.workers = ThreadPoolExecutor(max_workers=min(self.prepare_ahead, PREPARE_WORKERS)) \
            if self.prepare_ahead else None

    def close(self):
        if self.workers:
            self.workers.shutdown(wait=True)
        self.pool.close()

    def submit(self, rid, system, nsteps, arrival):
        self.queue.append([rid, arrival, system, nsteps, None])
        self._prefetch()

    def _prefetch(self):
        if not self.workers:
            return
        for q in list(self.queue)[:self.prepare_ahead]:
            if q[4] is None:
                q[4] = self.workers.submit(abi.prepare_job, q[2])

    def admit(self, now):
        """Admit waiting requests in arrival order while they fit."""
        while self.queue:
            rid, arrival, system, nsteps, fut = self.queue[0]
            if fut is None:
                job = self.pool.try_add_job(system, nsteps)
            else:
                if self.live and not fut.done():
                    break          # keep stepping; place it once it is laid out
                blob = fut.result()
                job = self.pool.add_prepared(blob, system, nsteps)
                if job is not None:
                    blob.free()
            if job is None:
                if not self.live:
                    raise RuntimeError(f"request {rid} does not fit an empty pool")
                break
            self.queue.popleft()
            self.live[job] = (rid, arrival, now, system, nsteps)
            self._prefetch()
        self.peak_live = max(self.peak_live, len(self.live))

    def advance(self):
        """One pd_step call over the resident jobs; returns the steps taken.
        A job the library stops during the call (PD_E_OVERFLOW: contact
        candidates beyond what a list can hold) does not stop the others: the
        call completes for them and harvest() retires the stopped one."""
        rem = [r for r, _ in (self.pool.status(j) for j in self.live) if r > 0]
        if not rem:
            return 0
        n = int(min(min(rem), self.chunk))
        try:
            self.pool.step(n)
        except abi.PDError as e:
            if e.code != abi.PD_E_OVERFLOW:
                raise
        self.steps_run += n
        self.device_ms += self.pool.last_step_ms()
        return n

    def harvest(self, now_fn):
        for job in list(self.live):
            left, why = self.pool.status(job)
            if left:
                continue
            rid, arrival, admitted, system, nsteps = self.live.pop(job)
            res = None
            if self.digests and why is None:
                res = read_results(lambda p: self.pool.state(job, p),
                                   lambda p: self.pool.connectivity(job, p),
                                   lambda: self.pool.wall(job), len(system.particles))
            self.pool.remove_job(job)
            self.done.append(Done(rid, arrival, admitted, now_fn(), nsteps, system.nodes, error=why))
            self.results.append(res)

    def serve(self, requests, clock=time.perf_counter):
        """requests: [(arrival seconds from start, system, nsteps)] in arrival order.
        Runs until every request is done; returns the Done records."""
        t0 = clock()
        now = lambda: clock() - t0
        pending = deque(enumerate(requests))
        while pending or self.queue or self.live:
            t = now()
            while pending and pending[0][1][0] <= t:
                rid, (arr, sysm, n) = pending.popleft()
                self.submit(rid, sysm, n, arr)
            self.admit(t)
            if not self.live:
                if pending:
                    time.sleep(max(0.0, pending[0][1][0] - now()))
                continue
            self.advance()
            self.harvest(now)
        # the digests verify the results; they are formed once serving is over
        for d, res in zip(self.done, self.results):
            if res is not None:
                d.digest = digest_arrays(res)
        self.results = []
        return self.done


def latency_stats(lat):
    lat = np.asarray(lat, dtype=float)
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
