"""Job pool: independent systems stepped together give the results they give alone.

A pool holds ragged jobs (different node counts, materials, speeds, walls,
switches and step budgets).  Each job is harvested when its budget is used
up and compared, bit for bit, with the same system run alone through the
plain interface (itself pinned to the oracle by the parity tests) and, for a
few jobs, with the oracle directly.  Tolerance: 0 ulp.
"""
import numpy as np
import pytest

from helpers import FIELDS, assert_bitwise
from peridem_amd import abi, gen

pytestmark = pytest.mark.gpu


def small_jobs(n, seed, walls_every=3):
    """Ragged two-sphere impacts; every walls_every-th one inside moving walls
    with reaction, some with a ramped external force."""
    out = gen.impact_batch(n, seed=seed, nodes_range=(60, 700), speed_range=(20.0, 120.0))
    for k, s in enumerate(out):
        if k % walls_every == 0:
            pos = np.concatenate([p.pos for p in s.particles])
            lo, hi = pos.min(0) - 0.3e-4, pos.max(0) + 0.3e-4
            s.allow_wall = 1
            s.compute_wall_reaction = 1
            s.wall_lrtb = tuple(lo) + tuple(hi)
            s.wall_speed = (0.0, 2.0, 0.0, -15.0, 0.0, -1.0)
        if k % 4 == 1:
            for p in s.particles:
                p.ext = np.zeros((p.n, 3))
                p.ext[:, 1] = -3e9
            s.gradient_extforce = 1
            s.extforce_maxstep = 17
        if k % 5 == 2:
            s.allow_friction = 0
        if k % 5 == 3:
            s.dt = 7e-9
    return out


def alone(system, nsteps):
    e = abi.gpu_engine(system)
    try:
        e.step(nsteps)
        st = [e.state(p) for p in range(len(system.particles))]
        cn = [e.connectivity(p) for p in range(len(system.particles))]
        wl = e.wall()
        return st, cn, wl
    finally:
        e.close()


def compare_job(pool, job, ref, label):
    st, cn, (lrtb, reac) = ref
    for p in range(len(st)):
        got = pool.state(job, p)
        for k in FIELDS:
            assert_bitwise(got[k], st[p][k], f"{label} particle {p} {k}")
        c = pool.connectivity(job, p)
        assert c.shape == cn[p].shape and np.array_equal(c, cn[p]), f"{label} particle {p} bonds"
    gl, gr = pool.wall(job)
    assert_bitwise(gl, lrtb, f"{label} wall planes")
    assert_bitwise(gr, reac, f"{label} wall reaction")


def pool_for(systems, jobs_cap=None, slack=1.0):
    fp = [abi.footprint(s) for s in systems]
    return abi.Pool(int(sum(f[0] for f in fp) * slack), int(sum(f[1] for f in fp) * slack),
                    sum(len(s.particles) for s in systems), jobs_cap or len(systems))


def test_footprint_counts_slices():
    s = small_jobs(1, seed=3)[0]
    ns, ne, nb = abi.footprint(s)
    assert ns % 64 == 0 and ns >= s.nodes and ns < s.nodes + 64
    assert ne % 64 == 0 and ne >= s.bonds
    assert nb > 0


def test_ragged_pool_matches_alone(oracle_lib):
    systems = small_jobs(7, seed=11)
    budgets = [40, 64, 25, 80, 33, 57, 12]
    print(f"[pool] 7 jobs nodes={[s.nodes for s in systems]} budgets={budgets}", flush=True)
    refs = [alone(s, n) for s, n in zip(systems, budgets)]
    # two of them against the oracle too
    for k in (0, 1):
        o = oracle_lib.engine(systems[k])
        o.step(budgets[k])
        for p in range(len(systems[k].particles)):
            so = o.state(p)
            for f in FIELDS:
                assert_bitwise(refs[k][0][p][f], so[f], f"alone job {k} vs oracle particle {p} {f}")
        o.close()
    pool = pool_for(systems)
    try:
        jobs = [pool.add_job(s, n) for s, n in zip(systems, budgets)]
        done = set()
        calls = 0
        while len(done) < len(jobs):
            pool.step(16 if calls % 2 == 0 else 5)   # graph replays and single launches
            calls += 1
            for k, j in enumerate(jobs):
                if k not in done and pool.remaining(j) == 0:
                    compare_job(pool, j, refs[k], f"job {k} (budget {budgets[k]})")
                    done.add(k)
        # frozen jobs stay as they were
        pool.step(9)
        for k, j in enumerate(jobs):
            compare_job(pool, j, refs[k], f"frozen job {k}")
    finally:
        pool.close()


def test_continuous_admission_reuses_freed_space():
    systems = small_jobs(9, seed=23)
    budgets = [30, 45, 18, 60, 27, 36, 50, 21, 40]
    refs = [alone(s, n) for s, n in zip(systems, budgets)]
    # room for about four jobs at a time
    pool = pool_for(systems[:4], jobs_cap=4, slack=1.6)
    try:
        queue = list(range(len(systems)))
        live = {}
        admitted_late = 0
        calls = 0
        while queue or live:
            while queue:
                j = pool.try_add_job(systems[queue[0]], budgets[queue[0]])
                if j is None:
                    break
                if calls:
                    admitted_late += 1
                live[j] = queue.pop(0)
            assert live, "a job must fit an empty pool"
            pool.step(8 if calls % 3 else 3)
            calls += 1
            for j in list(live):
                if pool.remaining(j) == 0:
                    k = live.pop(j)
                    compare_job(pool, j, refs[k], f"job {k}")
                    pool.remove_job(j)
        assert admitted_late > 0
    finally:
        pool.close()


def test_pool_rejects_what_does_not_fit():
    systems = small_jobs(2, seed=5)
    pool = pool_for(systems[:1])
    try:
        assert pool.try_add_job(systems[0], 10) is not None
        assert pool.try_add_job(systems[1], 10) is None
        pool.step(1)
        with pytest.raises(abi.PDError):
            pool.remove_job(7)
    finally:
        pool.close()


def test_graph_replay_matches_eager_launches():
    """Runs of 8 steps replay a captured graph; the same runs launched kernel
    by kernel give the same bits, over several replays with reads between."""
    s = small_jobs(4, seed=31)[0]
    a, b = abi.gpu_engine(s), abi.gpu_engine(s)
    try:
        b.set_graphs(False)
        for call, n in enumerate((8, 24, 16, 3, 40)):
            a.step(n)
            b.step(n)
            for p in range(len(s.particles)):
                sa, sb = a.state(p), b.state(p)
                for k in FIELDS:
                    assert_bitwise(sa[k], sb[k], f"call {call} particle {p} {k}")
            assert_bitwise(a.wall()[1], b.wall()[1], f"call {call} reaction")
    finally:
        a.close()
        b.close()
