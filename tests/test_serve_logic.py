"""Host logic of the continuous-batching server, on the CPU: a stand-in pool
with the same interface records what the server asks of it."""
import itertools

import numpy as np
import pytest

from peridem_amd import serve


class FakePool:
    """Capacity in abstract units; every job costs its system's `size`."""

    def __init__(self, capacity):
        self.capacity, self.used = capacity, 0
        self.jobs, self.calls = {}, []
        self.ids = itertools.count()

    def try_add_job(self, system, nsteps):
        if self.used + system.size > self.capacity:
            return None
        j = next(self.ids)
        self.jobs[j] = [system, nsteps]
        self.used += system.size
        self.calls.append(("add", system.name, nsteps))
        return j

    def remaining(self, j):
        return self.jobs[j][1]

    def status(self, j):
        return (0, self.jobs[j][2]) if len(self.jobs[j]) > 2 else (self.jobs[j][1], None)

    def step(self, n):
        assert n > 0 and all(v[1] >= n for v in self.jobs.values() if v[1]), "stepped past a budget"
        stopped = []
        for j, v in self.jobs.items():
            if getattr(v[0], "stop_after", None) is not None and len(v) == 2:
                # the library stops this job inside the call (PD_E_OVERFLOW)
                v[1] = 0
                v.append(f"job {j} stopped")
                stopped.append(j)
            v[1] = max(0, v[1] - n)
        self.calls.append(("step", n))
        if stopped:
            raise serve.abi.PDError("pd_step: PD_E_OVERFLOW: job stopped", serve.abi.PD_E_OVERFLOW)

    def last_step_ms(self):
        return 1.0

    def remove_job(self, j):
        self.used -= self.jobs.pop(j)[0].size
        self.calls.append(("remove", j))

    def close(self):
        pass


class Sys:
    def __init__(self, name, size):
        self.name, self.size, self.nodes = name, size, size
        self.particles = [None]


def make_server(capacity, chunk):
    srv = serve.Server.__new__(serve.Server)
    srv.max_nodes = srv.max_bond_entries = srv.max_particles = srv.max_jobs = 0
    srv.chunk, srv.digests, srv.prepare_ahead = chunk, False, 0
    srv.queue, srv.live, srv.done, srv.results = serve.deque(), {}, [], []
    srv.pool = FakePool(capacity)
    srv.steps_run, srv.device_ms, srv.peak_live, srv.workers = 0, 0.0, 0, None
    return srv


def test_admission_in_arrival_order_by_footprint():
    srv = make_server(capacity=10, chunk=50)
    reqs = [(0.0, Sys("a", 6), 30), (0.0, Sys("b", 5), 10), (0.0, Sys("c", 3), 20), (0.0, Sys("d", 4), 5)]
    done = srv.serve(reqs, clock=iter(np.arange(0, 100, 0.01)).__next__)
    adds = [c[1] for c in srv.pool.calls if c[0] == "add"]
    assert adds == ["a", "b", "c", "d"]               # arrival order, never overtaken
    # b (5) does not fit next to a (6): it waits for a's removal
    first_remove = next(i for i, c in enumerate(srv.pool.calls) if c[0] == "remove")
    assert srv.pool.calls.index(("add", "b", 10)) > first_remove
    assert sorted(d.rid for d in done) == [0, 1, 2, 3]


def test_steps_stop_at_the_earliest_budget_and_chunk():
    srv = make_server(capacity=100, chunk=16)
    reqs = [(0.0, Sys("a", 1), 40), (0.0, Sys("b", 1), 7)]
    srv.serve(reqs, clock=iter(np.arange(0, 100, 0.01)).__next__)
    steps = [c[1] for c in srv.pool.calls if c[0] == "step"]
    assert steps[0] == 7                              # b's budget ends the first call
    assert all(s <= 16 for s in steps) and sum(steps) == 40
    assert srv.steps_run == 40 and srv.peak_live == 2


def test_job_larger_than_the_pool_is_an_error():
    srv = make_server(capacity=3, chunk=8)
    with pytest.raises(RuntimeError, match="does not fit an empty pool"):
        srv.serve([(0.0, Sys("big", 5), 4)], clock=iter(np.arange(0, 10, 0.01)).__next__)


def test_latency_bookkeeping_and_fifo_reference():
    srv = make_server(capacity=100, chunk=8)
    clock = iter(np.arange(0, 1000, 0.5)).__next__
    done = srv.serve([(0.0, Sys("a", 1), 8), (0.0, Sys("b", 1), 16)], clock=clock)
    for d in done:
        assert d.finished >= d.admitted >= d.arrival and d.latency == d.finished - d.arrival
    assert serve.fifo_latency([0.0, 0.0, 5.0], [2.0, 3.0, 1.0]) == [2.0, 5.0, 1.0]
    st = serve.latency_stats([1.0, 2.0, 3.0, 4.0])
    assert st["p50_s"] == 2.5 and st["max_s"] == 4.0


def test_a_stopped_job_is_retired_and_serving_goes_on():
    """A job the library stops (PD_E_OVERFLOW from pd_step) is retired with
    its reason and no results; the other jobs run to their budgets."""
    srv = make_server(capacity=100, chunk=8)
    bad = Sys("bad", 1)
    bad.stop_after = 0
    done = srv.serve([(0.0, Sys("a", 1), 24), (0.0, bad, 30), (0.0, Sys("c", 1), 12)],
                     clock=iter(np.arange(0, 1000, 0.5)).__next__)
    by = {d.rid: d for d in done}
    assert sorted(by) == [0, 1, 2]
    assert by[1].error == "job 1 stopped" and by[0].error is None and by[2].error is None
    assert sum(c[1] for c in srv.pool.calls if c[0] == "step") == 24
