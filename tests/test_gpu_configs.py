"""The batched and served workloads at their own shapes.

Batched: 256 independent two-sphere impacts with node counts per sphere
log-uniform on [500, 20000], speeds U(5, 50) m/s and Young's modulus
U(20, 100) GPa (2.95M nodes, 272M directed bonds), stepped 30 steps in one
pool.  Served: a Poisson stream of aggregate-breakage jobs, each 2-64
seeded spheres and cubes of 2k-8k nodes inside walls with self-contact, with
their step budgets scaled by 1/1000 (5-50 steps), through continuous
batching.  Every job's results (disp, CurrPos, vel, acc, force, intact
bonds, wall planes and reaction of every particle) must equal, bit for bit,
the same system run alone through the plain interface, and sampled jobs the
oracle's -- each job is one run_timeloop (src/particle/timeloop.h:231-233).
Tolerance: 0 ulp.
"""
import numpy as np
import pytest

from peridem_amd import gen, serve

pytestmark = pytest.mark.gpu


def oracle_digest(oracle_lib, system, nsteps):
    o = oracle_lib.engine(system)
    try:
        o.step(nsteps)
        return serve.digest_state(o.state, o.connectivity, o.wall, len(system.particles))
    finally:
        o.close()


def test_batched_impacts_full_size_equal_alone_and_oracle(oracle_lib):
    systems = gen.impact_batch(256)
    nodes = [s.nodes for s in systems]
    print(f"[configs] batched: 256 impacts, {sum(nodes)} nodes, {min(nodes)}..{max(nodes)} per job", flush=True)
    wall_b, _, batch_d, _ = serve.run_batch(systems, 30, digests=True)
    print(f"[configs] batch done in {wall_b:.2f} s", flush=True)
    _, alone_d = serve.run_sequential(systems, 30, digests=True)
    same = sum(a == b for a, b in zip(batch_d, alone_d))
    assert same == 256, f"{256 - same} of 256 batched jobs differ from their solo runs"
    # the smallest job and the median one against the oracle
    order = np.argsort(nodes)
    for k in (int(order[0]), int(order[len(order) // 2])):
        print(f"[configs] oracle check of job {k} ({nodes[k]} nodes)", flush=True)
        assert oracle_digest(oracle_lib, systems[k], 30) == batch_d[k], f"job {k} differs from the oracle"


def test_served_aggregate_stream_equal_alone_and_oracle(oracle_lib):
    systems = [gen.aggregate_job(k) for k in range(8)]
    budgets = [max(5, s.meta["steps"] // 1000) for s in systems]
    arrivals = serve.poisson_arrivals(len(systems), rate=50.0, seed=5)
    print(f"[configs] served: particles={[s.meta['nparticles'] for s in systems]} "
          f"nodes={[s.nodes for s in systems]} budgets={budgets}", flush=True)
    _, alone_d = serve.run_sequential(systems, budgets, digests=True)
    cap = serve.pool_capacity(sorted(systems, key=lambda s: s.nodes)[-3:], slack=1.1)
    srv = serve.Server(chunk=16, **cap)
    try:
        done = srv.serve(list(zip(arrivals, systems, budgets)))
    finally:
        srv.close()
    assert sorted(d.rid for d in done) == list(range(len(systems)))
    assert all(d.error is None for d in done)
    assert srv.peak_live > 1, "requests never overlapped on the device"
    for d in done:
        assert d.digest == alone_d[d.rid], f"request {d.rid} differs from its solo run"
    k = int(np.argmin([s.nodes for s in systems]))
    print(f"[configs] oracle check of request {k} ({systems[k].nodes} nodes, {budgets[k]} steps)", flush=True)
    served = {d.rid: d.digest for d in done}
    assert oracle_digest(oracle_lib, systems[k], budgets[k]) == served[k]
