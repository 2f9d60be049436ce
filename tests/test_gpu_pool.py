"""Job pool: independent systems stepped together give the results they give alone.

A pool holds ragged jobs (different node counts, materials, speeds, walls,
switches and step budgets).  Each job is harvested when its budget is used
up and compared, bit for bit, with the same system run alone through the
plain interface (itself pinned to the oracle by the parity tests) and, for a
few jobs, with the oracle directly.  Tolerance: 0 ulp.
"""
import numpy as np
import pytest

from helpers import FIELDS, assert_bitwise, small_jobs
from peridem_amd import abi, gen

pytestmark = pytest.mark.gpu


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
    # jobs occupy whole 256-node blocks (4 slices of 64)
    assert ns % 256 == 0 and ns >= s.nodes and ns < s.nodes + 256
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


def test_reused_particle_slots_keep_pair_roles():
    """A job admitted into freed particle slots gets them in another order
    than its particle order; contact roles (which particle of a pair is
    P_C, whose box is extended by R_c) must still follow the job's particle
    order.  Two colliding sphere pairs with walls and reaction."""
    from helpers import small_impact
    a, b, c = (small_impact(seed=k, R_h=3.0) for k in (1, 2, 3))
    nsteps = 140
    ref = alone(c, nsteps)
    pool = pool_for([a, b], slack=1.6)
    try:
        ja = pool.add_job(a, 1000)
        pool.add_job(b, 1000)
        pool.step(10)
        pool.remove_job(ja)
        jc = pool.add_job(c, nsteps)      # takes a's slots, in reverse order
        while pool.remaining(jc):
            pool.step(min(pool.remaining(jc), 24))
        compare_job(pool, jc, ref, "job in reused slots")
    finally:
        pool.close()


def test_profiled_graph_replay_keeps_results_and_times_every_step():
    """With profiling on, steps launch one by one with phase marks that are
    read some steps later; the bits equal an unprofiled (graph) run, every
    step of every phase is counted once, calls shorter and longer than the
    read lag included, and the phase times add up to no more than the
    device time of the calls."""
    s = small_jobs(4, seed=37)[0]
    a, b = abi.gpu_engine(s), abi.gpu_engine(s)
    try:
        a.set_profiling(True)
        dev_ms, total = 0.0, 0
        for n in (8, 24, 5, 16):
            a.step(n)
            b.step(n)
            dev_ms += a.last_step_ms()
            total += n
        pr = a.profile()
        for name, ph in pr.items():
            assert ph["launches"] == total, (name, ph)
            assert ph["ms"] >= 0.0, (name, ph)
        summed = sum(ph["ms"] for ph in pr.values())
        assert 0.0 < summed <= dev_ms * 1.02 + 0.05, (summed, dev_ms)
        assert pr["bond"]["ms"] > 0.0
        for p in range(len(s.particles)):
            sa, sb = a.state(p), b.state(p)
            for k in FIELDS:
                assert_bitwise(sa[k], sb[k], f"particle {p} {k}")
        a.set_profiling(False)
        a.step(8)
        b.step(8)
        for p in range(len(s.particles)):
            assert_bitwise(a.state(p)["disp"], b.state(p)["disp"], f"after profiling, particle {p}")
    finally:
        a.close()
        b.close()


def test_admitting_a_job_with_overflowing_block_tables_switches_shape():
    """A pool steps a job with the staged bond kernel; a job whose 256-node
    blocks reference more nodes than the staged kernel's tables hold (a
    horizon of 5h) is admitted mid-run, so launches fall back to the
    unstaged kernel until it leaves, then return to the staged one.  Both
    jobs equal their solo runs bit for bit."""
    from helpers import small_impact
    a = small_impact(seed=23, R_h=3.0)
    rng = np.random.default_rng(29)
    wide = gen.make_particle(gen.lattice_cube(8), 1e-4, np.zeros(3), rng, horizon=5.0)
    wide.vel = rng.uniform(-2.0, 2.0, size=(wide.n, 3))
    b = gen.System(particles=[wide], dt=1e-8)
    shapes = abi.bond_shapes()
    staged = next(k for k, sh in enumerate(shapes) if sh[2])
    ref_a = alone(a, 60)
    ref_b = alone(b, 20)
    pool = pool_for([a, b], slack=1.2)
    try:
        ja = pool.add_job(a, 60)
        pool.set_bond_shape(staged)
        pool.step(20)
        assert pool.bond_shape() == staged
        jb = pool.add_job(b, 20)
        pool.step(20)                       # both jobs: the unstaged default runs
        assert not shapes[pool.bond_shape()][2]
        compare_job(pool, jb, ref_b, "wide job")
        pool.remove_job(jb)
        assert pool.bond_shape() == staged  # the staged shape is usable again
        pool.step(20)
        compare_job(pool, ja, ref_a, "job across the switches")
    finally:
        pool.close()


def test_bond_only_profiling_counts_the_bond_phase_only():
    s = small_jobs(4, seed=41)[0]
    a, b = abi.gpu_engine(s), abi.gpu_engine(s)
    try:
        a.set_profiling(abi.PD_PROFILE_BOND)
        for n in (3, 20):
            a.step(n)
            b.step(n)
        pr = a.profile()
        assert pr["bond"]["launches"] == 23 and pr["bond"]["ms"] > 0.0
        for name in ("position", "contact_list", "contact", "update"):
            assert pr[name]["launches"] == 0, (name, pr[name])
        a.set_profiling(0)
        for p in range(len(s.particles)):
            assert_bitwise(a.state(p)["force"], b.state(p)["force"], f"particle {p} force")
    finally:
        a.close()
        b.close()
