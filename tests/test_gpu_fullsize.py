"""The headline workload at its full size: 125 spheres, 1,003,125 nodes,
92,664,286 directed bonds (the bench's gravity-settling pool).

The oracle cannot step the whole pool in seconds, so the full-size checks are
properties that hold at any size, plus bitwise parity of a sample:

* rigid motion: with one velocity per sphere and no gravity every bond keeps
  exactly zero stretch, so every node's force is exactly zero;
* linear momentum: bond forces come in antisymmetric pairs, so
  sum_i V_i f_i = 0 per sphere up to rounding (no walls, no gravity, spheres
  apart);
* a sample: two of the 125 spheres, stepped inside the full pool, equal the
  oracle stepping those two spheres alone, bit for bit (they are far from
  the others and from the walls for these few steps).

The crushing aggregate (3.56M nodes) is checked in dense contact by teacher
forcing: the oracle is built from the GPU's state and both take one step.
"""
import time

import numpy as np
import pytest

from helpers import FIELDS, assert_bitwise
from peridem_amd import abi, gen

pytestmark = pytest.mark.gpu

NSPHERES = 125
# checkpoints of the crushing aggregate's teacher-forced steps: the top wall
# has driven the top layers into contact with each other (A) and the packing
# is in contact through most of its height (B); at least this many particle
# pairs pass the box test in the checked step
DENSE_STEPS_A, DENSE_STEPS_B = 20000, 60000
MIN_DENSE_PAIRS = 50


def full_pool(**kw):
    s = gen.gravity_settling(nspheres=NSPHERES, **kw)
    print(f"[fullsize] {len(s.particles)} spheres, {s.nodes} nodes, {s.bonds} directed bonds", flush=True)
    assert s.nodes == 1003125 and s.bonds == 92664286
    return s


def without_gravity(s):
    for p in s.particles:
        p.ext = np.zeros((p.n, 3))
    return s


def test_rigid_motion_leaves_every_bond_unstretched():
    s = without_gravity(full_pool(sphere_speed=0.5, node_speed=0.0))
    s.allow_wall = 0
    e = abi.gpu_engine(s)
    try:
        e.step(10)
        c = e.counters()
        assert c["alive_bonds"] == s.bonds
        for p in range(NSPHERES):
            st = e.state(p)
            f = st["force"]
            assert not np.any(f.view(np.uint64)), f"sphere {p}: force not exactly zero"
            # every node of a sphere moved by the same displacement
            d = st["disp"]
            assert np.all(d == d[0]), f"sphere {p}: displacement not rigid"
    finally:
        e.close()


def test_bond_forces_conserve_linear_momentum():
    s = without_gravity(full_pool(sphere_speed=0.5, node_speed=0.01))
    s.allow_wall = 0
    e = abi.gpu_engine(s)
    try:
        e.step(20)
        worst = 0.0
        for p in range(NSPHERES):
            f = e.state(p)["force"]
            vol = s.particles[p].vol[:, None]
            vf = vol * f
            net = np.abs(vf.sum(axis=0)).max()
            scale = np.abs(vf).sum(axis=0).max()
            assert scale > 0.0
            worst = max(worst, net / scale)
        print(f"[fullsize] worst |sum V f| / sum |V f| = {worst:.3e}", flush=True)
        assert worst < 1e-12
    finally:
        e.close()


def test_two_sphere_sample_matches_oracle_bitwise(oracle_lib):
    s = full_pool(sphere_speed=0.5, node_speed=0.01)
    sample = gen.System(**{**s.__dict__, "particles": s.particles[:2]})
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(sample)
    try:
        done = 0
        for c in (1, 5, 12):
            g.step(c - done)
            o.step(c - done)
            done = c
            for p in range(2):
                sg, so = g.state(p), o.state(p)
                for k in FIELDS:
                    assert_bitwise(sg[k], so[k], f"step {c} sphere {p} {k}")
                cg, co = g.connectivity(p), o.connectivity(p)
                assert np.array_equal(cg, co), f"step {c} sphere {p} connectivity"
        assert_bitwise(g.wall()[0], o.wall()[0], "wall planes")
    finally:
        g.close()
        o.close()


def test_crushing_aggregate_full_size_sample_matches_oracle(oracle_lib):
    """The crushing-aggregate configuration at full size (512 cubes, spheres
    and convex polyhedra of 2k-12k nodes, 3.56M nodes, 324M directed bonds,
    self-contact, walls with the top one moving, wall reaction on), every
    node with a seeded velocity so every bond carries stretch.  The particles
    are placed with their bounding spheres at least 1.5 R_c apart and clear
    of the walls, so particle pairs pass the box test (the boxes are extended
    by R_c) but for these few steps (nodes move < 1e-7 m) no node comes within
    R_c of another particle's node or of a wall: two of them, stepped inside
    the full pool, must equal the oracle stepping those two alone with the
    same walls, bit for bit."""
    s = gen.crushing_aggregate(512)
    rng = np.random.default_rng(17)
    for p in s.particles:
        p.vel = rng.uniform(-1.0, 1.0, size=(p.n, 3))
    print(f"[fullsize] crushing aggregate: {len(s.particles)} particles, {s.nodes} nodes, "
          f"{s.bonds} directed bonds", flush=True)
    assert s.nodes == 3563127 and s.bonds == 324108448
    sample = gen.System(**{**s.__dict__, "particles": s.particles[:2]})
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(sample)
    try:
        done = 0
        for c in (1, 4, 10):
            g.step(c - done)
            o.step(c - done)
            done = c
            for p in range(2):
                sg, so = g.state(p), o.state(p)
                for k in FIELDS:
                    assert_bitwise(sg[k], so[k], f"step {c} particle {p} {k}")
                assert np.array_equal(g.connectivity(p), o.connectivity(p)), f"step {c} particle {p} bonds"
            lg, rg = g.wall()
            lo, ro = o.wall()
            assert_bitwise(lg, lo, f"step {c} wall planes")
            assert not np.any(rg), "no node touches a wall yet"
        cnt = g.counters()
        assert cnt["alive_bonds"] == s.bonds
        print(f"[fullsize] {cnt['pair_computations']} particle pairs passed the box test in 10 steps "
              f"(none within R_c)", flush=True)
    finally:
        g.close()
        o.close()


def state_as_system(s, g, first_step):
    """The system as the GPU holds it after some steps -- every particle's
    disp, vel, acc and intact bonds (Connectivity as NbdArr2conn emits it,
    timeloop.h:20-44) and the wall planes: what a tc_/wall_ checkpoint holds
    (timeloop.h:262-305), with the step counter continued."""
    parts = []
    for p, P in enumerate(s.particles):
        st = g.state(p)
        Q = gen.Particle(**P.__dict__)
        Q.disp, Q.vel, Q.acc = st["disp"], st["vel"], st["acc"]
        Q.conn = g.connectivity(p)
        parts.append(Q)
    lrtb, _ = g.wall()
    return gen.System(**{**s.__dict__, "particles": parts, "wall_lrtb": tuple(lrtb),
                         "first_step": first_step})


@pytest.mark.parametrize("checkpoints", [(DENSE_STEPS_A, DENSE_STEPS_B)])
def test_crushing_aggregate_dense_contact_teacher_forced(oracle_lib, checkpoints):
    """The full crushing aggregate (512 particles, 3.56M nodes, 324M directed
    bonds) stepped on the GPU while the top wall drives the packing into
    contact; at each checkpoint the oracle is built from the GPU's state and
    both take one step, which must agree bit for bit on every node, bond,
    wall plane, wall reaction and on the number of particle pairs in contact
    (contact.h:746-942, 1060-1124, timeloop.h:636-651).  The oracle prunes
    its contact candidates with a cell grid, which leaves its sums bitwise
    those of its all-pairs loops (test_pruned_contact_is_bitwise_the_all_pairs_loop)."""
    s = gen.crushing_aggregate(512)
    g = abi.gpu_engine(s)
    done = 0
    try:
        for k in checkpoints:
            g.step(k - done)
            done = k
            t0 = time.perf_counter()
            st = state_as_system(s, g, first_step=done + 1)
            o = oracle_lib.engine(st, prune=True)
            pw0 = g.counters()["pair_computations"]
            print(f"[fullsize] teacher-forced step {done + 1}: oracle built in "
                  f"{time.perf_counter() - t0:.1f} s", flush=True)
            try:
                g.step(1)
                o.step(1)
                done += 1
                pairs = g.counters()["pair_computations"] - pw0
                assert pairs == o.counters()["pair_computations"], (pairs, o.counters())
                for p in range(len(s.particles)):
                    sg, so = g.state(p), o.state(p)
                    for f in FIELDS:
                        assert_bitwise(sg[f], so[f], f"step {done} particle {p} {f}")
                    assert np.array_equal(g.connectivity(p), o.connectivity(p)), f"step {done} particle {p} bonds"
                lg, rg = g.wall()
                lo, ro = o.wall()
                assert_bitwise(lg, lo, f"step {done} wall planes")
                assert_bitwise(rg, ro, f"step {done} wall reaction")
                assert np.any(rg[5] != 0), "the top wall touches the packing"
                print(f"[fullsize] step {done}: {pairs} particle pairs in contact, "
                      f"{s.bonds - g.counters()['alive_bonds']} directed bonds broken, bitwise equal",
                      flush=True)
                assert pairs >= MIN_DENSE_PAIRS
            finally:
                o.close()
    finally:
        g.close()
