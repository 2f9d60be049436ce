"""GPU parity: the HIP library against the CPU oracle, bitwise.

Every test builds one seeded system, runs it through libperidem_gpu.so (via
the C-ABI) and through the oracle, and compares disp, CurrPos, vel, acc,
force, the intact-bond Connectivity, wall planes, wall reaction and the step
counters bit for bit at several checkpoints.  Tolerance: 0 ulp (the library
evaluates every term in the reference's order with IEEE FP64 and no FMA).
"""
import numpy as np
import pytest

from helpers import assert_bitwise, compare_engines, degenerate_particles, mixed_system, small_impact
from peridem_amd import abi, gen

pytestmark = pytest.mark.gpu


def run_pair(system, checkpoints, oracle_lib, label, reaction=False, prune=False):
    """prune: the oracle's cell-grid contact candidates, bitwise its all-pairs
    loops (test_pruned_contact_is_bitwise_the_all_pairs_loop), for long runs."""
    print(f"[parity] {label}: nodes={system.nodes} bonds={system.bonds} checkpoints={checkpoints}",
          flush=True)
    g = abi.gpu_engine(system)
    o = oracle_lib.engine(system, prune=prune)
    done = 0
    try:
        for c in checkpoints:
            g.step(c - done)
            o.step(c - done)
            done = c
            rg, ro = compare_engines(g, o, system, f"{label} step {c}")
            if reaction:
                assert_bitwise(rg, ro, f"{label} step {c} wall reaction")
        return g.counters(), o.counters()
    finally:
        g.close()
        o.close()


def test_library_loads_and_reports_version():
    lib, _ = abi.load_gpu_lib()
    assert b"gfx950" in lib.pd_version()


def test_free_flight_two_spheres(oracle_lib):
    s = gen.two_sphere_impact(R=5e-4, h=1e-4, seed=3)
    run_pair(s, [1, 5], oracle_lib, "free flight")


def test_config0_two_sphere_impact(oracle_lib):
    # the reference's small config: two R=1e-3 spheres, h=1e-4, 20 m/s, fixed B;
    # contact begins near step 100; all 2000 steps of the configuration
    s = gen.two_sphere_impact()
    cg, co = run_pair(s, [1, 120, 300, 1000, 2000], oracle_lib, "config0", prune=True)
    assert cg["pair_computations"] > 0 and cg["steps"] == 2000


def test_small_impact_breaks_bonds_and_walls(oracle_lib):
    s = small_impact()
    cg, co = run_pair(s, [1, 3, 10, 25, 60], oracle_lib, "small impact", reaction=True)
    assert cg["alive_bonds"] < cg["total_bonds"], "impact should break bonds"
    assert cg["pair_computations"] > 0


def test_small_impact_no_self_contact_no_walls(oracle_lib):
    s = small_impact(seed=11, self_contact=0, walls=False, b_movable=0)
    run_pair(s, [2, 20, 50], oracle_lib, "impact without self-contact")


def test_mixed_flags_clamped_gradient_ext(oracle_lib):
    s = mixed_system()
    cg, _ = run_pair(s, [1, 2, 15, 40, 120], oracle_lib, "mixed", reaction=True)
    assert cg["pair_computations"] > 0


def test_time_switches(oracle_lib):
    s = small_impact(seed=21, self_contact=0)
    s.enable_fracture = 0
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        for e in (g, o):
            e.step(4)
            e.set_flag(abi.PD_FLAG_ENABLE_FRACTURE, 1)
            e.set_flag(abi.PD_FLAG_SELF_CONTACT, 1)
            e.step(6)
            e.set_wall_speed([0, 0, 0, 0, 0, 0])
            e.set_particle_flag(1, abi.PD_PFLAG_STOPPABLE, 0)
            e.set_snot(0, 1e-4)
            e.step(10)
        rg, ro = compare_engines(g, o, s, "switches")
        assert_bitwise(rg, ro, "switches reaction")
    finally:
        g.close()
        o.close()


def test_set_state_resume(oracle_lib):
    s = small_impact(seed=31, walls=False)
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        rng = np.random.default_rng(0)
        for p in range(2):
            n = s.particles[p].n
            d, v, a = (rng.uniform(-1e-6, 1e-6, (n, 3)), rng.uniform(-10, 10, (n, 3)),
                       rng.uniform(-1e4, 1e4, (n, 3)))
            g.set_state(p, d, v, a)
            o.set_state(p, d, v, a)
        g.step(8)
        o.step(8)
        compare_engines(g, o, s, "resume")
    finally:
        g.close()
        o.close()


def test_every_bond_shape_matches_oracle(oracle_lib):
    """Each instantiated bond-kernel shape gives the oracle's bits."""
    s = small_impact(seed=9)
    o = oracle_lib.engine(s)
    o.step(40)
    ref = [o.state(p) for p in range(len(s.particles))]
    refc = [o.connectivity(p) for p in range(len(s.particles))]
    o.close()
    for k, shape in enumerate(abi.bond_shapes()):
        print(f"[parity] bond shape {k} {shape}", flush=True)
        g = abi.gpu_engine(s)
        try:
            g.set_bond_shape(k)
            g.step(40)
            assert g.bond_shape() == k
            for p in range(len(s.particles)):
                st = g.state(p)
                for f in ("disp", "currpos", "vel", "acc", "force"):
                    assert_bitwise(st[f], ref[p][f], f"shape {k} particle {p} {f}")
                assert np.array_equal(g.connectivity(p), refc[p]), f"shape {k} particle {p} bonds"
        finally:
            g.close()


def test_small_pools_take_the_four_lane_shape(oracle_lib):
    """A pool below 2^15 resident nodes runs the bond kernel with four lanes
    per node slot (shape 14, pipelined = 3) unless a shape is forced; a larger
    one below the tuning size keeps the staged shape 10.  Through contact and
    breaks the four-lane run is the oracle's bit for bit."""
    quad = 14
    assert abi.bond_shape_info(quad)[1] == 3
    s = small_impact(seed=13)
    assert s.nodes < 1 << 15
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        for n in (1, 30, 60):
            print(f"[parity] four-lane default shape, {n} steps", flush=True)
            g.step(n)
            o.step(n)
            assert g.bond_shape() == quad
            compare_engines(g, o, s, "four-lane shape")
        c = g.counters()
        assert c["alive_bonds"] < c["total_bonds"] and c["pair_computations"] > 0
    finally:
        g.close()
        o.close()
    big = gen.aggregate_job(9)
    assert 1 << 15 <= big.nodes < 1 << 18
    g = abi.gpu_engine(big)
    try:
        g.step(1)
        assert g.bond_shape() == 10
    finally:
        g.close()


def test_wide_bond_shape_matches_oracle(oracle_lib):
    """The 64-bit-offset bond kernel of large pools, forced on a small one,
    gives the oracle's bits through a run that breaks bonds."""
    s = small_impact(seed=21)
    k = abi.wide_bond_shape()
    assert abi.bond_shape_info(k) == (4, 2, False)
    print(f"[parity] wide bond shape {k}", flush=True)
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        g.set_bond_shape(k)
        g.step(72)
        o.step(72)
        assert g.bond_shape() == k
        compare_engines(g, o, s, "wide shape")
        assert g.counters()["alive_bonds"] < g.counters()["total_bonds"], "no bond broke"
    finally:
        g.close()
        o.close()


def test_switching_between_staged_and_global_shapes_keeps_breaks(oracle_lib):
    """Both forms of a bond entry (pool index and position in the block's
    node table) carry the broken flag: alternating staged and unstaged shapes
    through a run that breaks bonds still gives the oracle's bits."""
    s = small_impact(seed=21)
    shapes = abi.bond_shapes()
    staged = [k for k, sh in enumerate(shapes) if sh[2]]
    plain = [k for k, sh in enumerate(shapes) if not sh[2]]
    assert staged and plain
    order = [staged[0], plain[0], staged[-1], plain[3], staged[1], plain[-1]]
    print(f"[parity] shape order {order}", flush=True)
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        done = 0
        for k in order:
            g.set_bond_shape(k)
            g.step(12)
            o.step(12)
            done += 12
            assert g.bond_shape() == k
            compare_engines(g, o, s, f"shape {k} to step {done}")
        assert g.counters()["alive_bonds"] < g.counters()["total_bonds"], "no bond broke"
    finally:
        g.close()
        o.close()


def test_staged_shapes_fall_back_when_block_tables_overflow(oracle_lib):
    """A horizon of 5h on a 17^3 cube makes 256-node blocks reference more
    nodes than the staged kernel's LDS table holds: a staged shape then falls
    back to the default unstaged one, the tuner skips staged shapes, and the
    result is still the oracle's."""
    rng = np.random.default_rng(5)
    p = gen.make_particle(gen.lattice_cube(8), 1e-4, np.zeros(3), rng, horizon=5.0)
    p.vel = rng.uniform(-2.0, 2.0, size=(p.n, 3))
    s = gen.System(particles=[p], dt=1e-8)
    print(f"[parity] overflowing block tables: nodes={s.nodes} bonds={s.bonds}", flush=True)
    shapes = abi.bond_shapes()
    staged = [k for k, sh in enumerate(shapes) if sh[2]]
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        g.set_bond_shape(staged[0])
        g.step(6)
        o.step(6)
        assert not shapes[g.bond_shape()][2]
        compare_engines(g, o, s, "overflowing block tables")
        best, ms = g.autotune(1)
        assert not shapes[best][2]
        assert all(ms[k] == -1 for k in staged) and all(ms[k] > 0 for k in range(len(shapes)) if k not in staged)
        assert ms[best] == min(ms[k] for k in range(len(shapes)) if ms[k] > 0)   # the fastest timed shape
        g.step(4)
        o.step(4)
        compare_engines(g, o, s, "after autotune")
    finally:
        g.close()
        o.close()


def test_autotune_leaves_state_untouched(oracle_lib):
    s = small_impact(seed=12)
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        g.step(15)
        o.step(15)
        before = [g.state(p) for p in range(len(s.particles))]
        best, ms = g.autotune(2)
        shapes = abi.bond_shapes()
        assert 0 <= best < len(shapes) and g.bond_shape() == best
        # staged shapes usable here: all of them timed, of the unstaged ones the
        # default (4) and, on a pool below 2^18 nodes, the four-lane shape
        assert all((ms[k] > 0) == (sh[2] or k == 4 or sh[1] == 3) for k, sh in enumerate(shapes))
        assert ms[best] == ms[ms > 0].min()          # the tuner keeps the fastest shape
        for p in range(len(s.particles)):
            after = g.state(p)
            for f in ("disp", "vel", "acc", "force"):
                assert_bitwise(after[f], before[p][f], f"autotune changed particle {p} {f}")
        g.step(25)
        o.step(25)
        compare_engines(g, o, s, "after autotune")
    finally:
        g.close()
        o.close()


def test_candidate_lists_rebuild_and_match_oracle(oracle_lib):
    """Contact candidates come from lists rebuilt when a node has moved half
    the skin: a fast impact rebuilds several times and reuses the lists in
    between, and still matches the oracle's all-pairs search bit for bit."""
    s = small_impact(seed=17, v=150.0)
    cg, co = run_pair(s, [40, 90, 160], oracle_lib, "candidate lists")
    assert 3 <= cg["cell_rebuilds"] < 160, cg["cell_rebuilds"]


def test_torque_matches_oracle(oracle_lib):
    """External torque about each particle's centroid (timeloop.h:601-633), on
    all four axes settings incl. a non-movable particle, switched on mid-run
    like set_torque_timestep."""
    s = mixed_system(seed=21)
    for p, (ax, tv) in zip(s.particles, [(2, 3.0e9), (0, -2.0e9), (1, 5.0e9), (2, 0.0)]):
        p.torque_axis, p.torque_val = ax, tv
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        for eng in (g, o):
            eng.step(12)
            eng.set_flag(abi.PD_FLAG_ENABLE_TORQUE, 1)
        compare_engines(g, o, s, "torque switch")
        for c in (20, 60):
            g.step(c)
            o.step(c)
            rg, ro = compare_engines(g, o, s, f"torque +{c}")
            assert_bitwise(rg, ro, f"torque +{c} reaction")
    finally:
        g.close()
        o.close()


def test_nl_bdry_only_matches_oracle(oracle_lib):
    """Contact and wall terms on boundary nodes only, accumulated in the order
    of each particle's boundary list (shuffled here), self-contact included."""
    s = small_impact(seed=23)
    s.nl_bdry_only = 1
    rng = np.random.default_rng(5)
    for p in s.particles:
        deg = np.bincount(p.conn.reshape(-1), minlength=p.n)
        surf = np.nonzero(deg < deg.max())[0].astype(np.uint32)
        p.bdry = rng.permutation(surf)
    run_pair(s, [30, 90, 150], oracle_lib, "nl_bdry_only", reaction=True)


@pytest.mark.parametrize("mean_x,angular", [(1, 0), (0, 1), (1, 1), (0, 0)])
def test_prescribed_velocity_matches_oracle(oracle_lib, mean_x, angular):
    """particle_w_fixed_velocity (timeloop.h:685-764): particle 0 (with clamped
    nodes) gets its mean x velocity and/or an angular velocity prescribed from
    a switch step on; state and applied_force_density bitwise."""
    s = mixed_system(seed=33)
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        for eng in (g, o):
            eng.set_velocity_constraint(0, fix_mean_x=mean_x, mean_x=-12.5, fix_angular=angular,
                                        angular_vel=4.0e4)
            eng.step(9)
            eng.set_flag(abi.PD_FLAG_VELOCITY_CONSTRAINT, 1)
        for c in (1, 16, 40):
            g.step(c)
            o.step(c)
            rg, ro = compare_engines(g, o, s, f"prescribed {mean_x}{angular} +{c}")
            assert_bitwise(g.applied_force(0), o.applied_force(0), f"applied force +{c}")
    finally:
        g.close()
        o.close()


def test_crushing_aggregate_force_displacement_trace(oracle_lib):
    """Config-2 shape at small scale: mixed cubes, spheres and convex polyhedra
    crushed by the moving top wall.  The force-displacement trace (top-wall
    position and reaction), the progressive loss of bonds and every particle's
    state match the oracle bit for bit."""
    s = gen.crushing_aggregate(nparticles=8, nodes_range=(150, 500), h=6e-5, top_speed=-60.0,
                               seed=13, dt=4e-9, packing=0.5)
    cg, co = run_pair(s, list(range(100, 801, 100)), oracle_lib, "crushing", reaction=True)
    broken = 1 - cg["alive_bonds"] / cg["total_bonds"]
    print(f"[parity] crushing: broken fraction {broken:.3f} after 800 steps", flush=True)
    assert 0.05 < broken < 0.9


def test_pair_activity_lags_one_step_like_the_oracle(oracle_lib):
    """A node crossing into R_c within one step: the pair test uses the
    previous step's boxes (timeloop.h:389-393,779), so contact starts one
    step late -- the library and the oracle agree bit for bit."""
    Rc, V = 9e-5, 1000.0
    A = gen.Particle(pos=np.zeros((1, 3)), vol=np.array([2e-12]), conn=np.zeros((0, 2), np.uint32),
                     delta=3e-4, rho=2000.0, cnot=3e25, snot=2e-3)
    A.vel = np.array([[V, 0.0, 0.0]])
    B = gen.Particle(pos=np.array([[Rc + 0.5 * V * 1e-8, 0.0, 0.0]]), vol=np.array([5e-12]),
                     conn=np.zeros((0, 2), np.uint32), delta=3e-4, rho=2440.0, cnot=3e25, snot=2e-3)
    s = gen.System(particles=[A, B], dt=1e-8, contact_rad=Rc, normal_stiffness=4.6e31,
                   friction_coefficient=0.3, damping_ratio=0.1, enable_fracture=1, self_contact=0)
    run_pair(s, (1, 2, 3), oracle_lib, "lagged pair boxes")


@pytest.mark.parametrize("first_reset", [1, 12])
def test_wheel_reset_matches_oracle(oracle_lib, first_reset):
    """reset_partzero_y (timeloop.h:496-535): the wheel is moved before step
    first_reset and again 20 steps later; wheel_rad = 2h drops it into the
    bulk below, so both resets are followed by wheel-bulk contact, self-contact,
    walls and bond breaks.  The bond kernel keeps the setup positions (xi)
    while contact and CurrPos use the moved ones: bitwise against the oracle
    after each reset and at later checkpoints."""
    from helpers import wheel_on_bulk
    s = wheel_on_bulk(wheel_vy=-3.0, walls=True, self_contact=1)
    wheel_rad = 2e-4
    print(f"[parity] wheel reset at {first_reset}: nodes={s.nodes} bonds={s.bonds}", flush=True)
    g, o = abi.gpu_engine(s), oracle_lib.engine(s)
    try:
        done = 0
        plan = [(first_reset - 1, True), (1, False), (19, True), (1, False), (30, False)]
        pairs0 = 0
        for n, reset in plan:
            if n:
                g.step(n)
                o.step(n)
                done += n
            if reset:
                g.reset_partzero_y(wheel_rad)
                o.reset_partzero_y(wheel_rad)
            else:
                rg, ro = compare_engines(g, o, s, f"wheel step {done}")
                assert_bitwise(rg, ro, f"wheel step {done} wall reaction")
            if done == first_reset:
                pairs0 = g.counters()["pair_computations"]
        assert g.counters()["pair_computations"] > pairs0, "the wheel lands in contact with the bulk"
    finally:
        g.close()
        o.close()


def test_degenerate_particles_match_oracle(oracle_lib):
    """Empty Connectivities (a single-node particle, three unbonded nodes, a
    bond-free cube whose nodes meet each other and the moving top wall) and an
    empty row inside a bonded sphere: slices of width zero, block tables of
    the block's own nodes only, self-contact between unbonded nodes, broken
    bonds, wall reaction -- bitwise against the oracle."""
    s = degenerate_particles()
    run_pair(s, [1, 100, 300, 600, 800], oracle_lib, "degenerate particles", reaction=True)
