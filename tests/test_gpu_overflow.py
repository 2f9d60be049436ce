"""Contact candidate lists longer than their capacity.

A node's list holds every node of its job within R_c + skin.  When a list
build finds more than the pool's capacity, the pool halts inside that step
(every later kernel of the call exits), the library goes back to the end of
the previous step (the displacement is double-buffered), grows the lists and
runs the step again, so the run is bit for bit the oracle's.  A job whose
lists would need more than 255 entries is stopped at the end of its last
complete step with PD_E_OVERFLOW naming it; the other jobs of the pool carry
on unaffected.  Tolerance: 0 ulp.
"""
import numpy as np
import pytest

from helpers import FIELDS, assert_bitwise, candidate_schedule, compare_engines, passing_spheres
from peridem_amd import abi, gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bdry_only", [0, 1])
def test_lists_grow_from_the_first_step(oracle_lib, bdry_only):
    """R_c = 2.2 h with self-contact puts ~70 nodes within the list radius of
    an interior node, more than the 48 a list first holds: the lists grow in
    the first step, in both list forms."""
    h = 1e-4
    s = gen.two_sphere_impact(R=4 * h, h=h, v=150.0, seed=31, gap_h=1.0, b_movable=1, self_contact=1)
    s.contact_rad = 2.2 * h
    s.normal_stiffness = gen.normal_stiffness(72e9, 0.25, s.contact_rad)
    s.nl_bdry_only = bdry_only
    if bdry_only:
        for p in s.particles:
            p.bdry = np.arange(p.n, dtype=np.uint32)
    print(f"[overflow] first step, nl_bdry_only={bdry_only}: nodes={s.nodes}", flush=True)
    g, o = abi.gpu_engine(s), oracle_lib.engine(s)
    try:
        g.step(12)
        o.step(12)
        compare_engines(g, o, s, "grown lists")
        c = g.counters()
        assert c["list_regrows"] >= 1 and c["list_capacity"] > 48, c
        assert c["pair_computations"] > 0
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("Rc_h", [1.8, 2.0])
def test_lists_grow_inside_a_graph_replay(oracle_lib, Rc_h):
    """Spheres passing through each other: the first build needs ~10 entries,
    a rebuild between steps 9 and 40 needs more than 48 (more than 80 too for
    R_c = 2h, a second growth), all inside one pd_step(40) call of five graph
    replays."""
    s = passing_spheres(Rc_h=Rc_h)
    sched = candidate_schedule(s, 40, oracle_lib)
    first_over = next(k for k, n in sched if n > 48)
    assert sched[0][1] <= 48 and 9 <= first_over <= 40, sched
    print(f"[overflow] mid-call, Rc={Rc_h}h: first rebuild over 48 at step {first_over}", flush=True)
    g, o = abi.gpu_engine(s), oracle_lib.engine(s)
    try:
        g.step(40)
        o.step(40)
        compare_engines(g, o, s, f"mid-call growth Rc={Rc_h}h")
        c = g.counters()
        need = max(n for _, n in sched)
        assert c["list_regrows"] >= 1 and c["list_capacity"] >= need, (c, sched)
        g.step(5)
        o.step(5)
        compare_engines(g, o, s, "after growth")
    finally:
        g.close()
        o.close()


def _oracle_run(oracle_lib, s, n):
    o = oracle_lib.engine(s)
    o.step(n)
    st = [o.state(p) for p in range(len(s.particles))]
    cn = [o.connectivity(p) for p in range(len(s.particles))]
    o.close()
    return st, cn


def _compare_job(pool, job, ref, label):
    st, cn = ref
    for p in range(len(st)):
        got = pool.state(job, p)
        for k in FIELDS:
            assert_bitwise(got[k], st[p][k], f"{label} particle {p} {k}")
        c = pool.connectivity(job, p)
        assert np.array_equal(c, cn[p]), f"{label} particle {p} bonds"


def test_growth_in_a_pool_keeps_every_job_exact(oracle_lib):
    """One job's lists grow mid-call; the other job of the pool (no overflow)
    and the growing one both end bit for bit on their oracle runs."""
    from helpers import small_impact
    grow = passing_spheres(Rc_h=2.0)
    calm = small_impact(seed=8)
    systems = [calm, grow]
    fp = [abi.footprint(x) for x in systems]
    pool = abi.Pool(sum(f[0] for f in fp), sum(f[1] for f in fp), 4, 2)
    try:
        jobs = [pool.add_job(x) for x in systems]
        pool.step(16)
        pool.step(19)
        assert pool.counters()["list_regrows"] >= 1
        for j, x in zip(jobs, systems):
            _compare_job(pool, j, _oracle_run(oracle_lib, x, 35), f"pool job {j}")
    finally:
        pool.close()


def test_a_job_beyond_the_largest_list_is_stopped_alone(oracle_lib):
    """R_c = 3.5 h with self-contact: ~390 nodes within the list radius of a
    central node, more than 255.  That job stops before its first step and
    pd_step reports it; the other job steps on and matches the oracle; the
    stopped job keeps its initial state and pd_job_remaining reports why."""
    from helpers import small_impact
    h = 1e-4
    rng = np.random.default_rng(41)
    big = gen.make_particle(gen.lattice_sphere(5.0), h, np.zeros(3), rng)
    bad = gen.System(particles=[big], dt=1e-8, contact_rad=3.5 * h,
                     normal_stiffness=gen.normal_stiffness(72e9, 0.25, 3.5 * h), self_contact=1,
                     enable_fracture=1)
    calm = small_impact(seed=9)
    systems = [calm, bad]
    fp = [abi.footprint(x) for x in systems]
    pool = abi.Pool(sum(f[0] for f in fp), sum(f[1] for f in fp), 3, 2)
    try:
        jc, jb = pool.add_job(calm), pool.add_job(bad)
        with pytest.raises(abi.PDError, match=f"job {jb} stopped"):
            pool.step(12)
        with pytest.raises(abi.PDError, match="contact candidates"):
            pool.remaining(jb)
        pool.step(6)                                  # the healthy job carries on; no error now
        _compare_job(pool, jc, _oracle_run(oracle_lib, calm, 18), "healthy job")
        st = pool.state(jb, 0)
        assert np.all(st["disp"] == 0) and np.all(st["vel"] == 0)
    finally:
        pool.close()


def test_plain_run_beyond_the_largest_list_keeps_reporting():
    h = 1e-4
    rng = np.random.default_rng(42)
    big = gen.make_particle(gen.lattice_sphere(5.0), h, np.zeros(3), rng)
    s = gen.System(particles=[big], dt=1e-8, contact_rad=3.5 * h,
                   normal_stiffness=gen.normal_stiffness(72e9, 0.25, 3.5 * h), self_contact=1,
                   enable_fracture=1)
    g = abi.gpu_engine(s)
    try:
        for _ in range(2):
            with pytest.raises(abi.PDError, match="contact candidates"):
                g.step(3)
        assert g.counters()["steps"] == 0
    finally:
        g.close()
