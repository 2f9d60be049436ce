"""Shared helpers of the parity tests: small seeded systems and comparisons."""
import numpy as np

from peridem_amd import gen

FIELDS = ("disp", "currpos", "vel", "acc", "force")


def assert_bitwise(a, b, what):
    """Bitwise equality of two float arrays (NaNs compare equal by bit pattern)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    ua, ub = a.view(np.uint64), b.view(np.uint64)
    bad = np.nonzero(ua != ub)
    if bad[0].size:
        i = tuple(x[0] for x in bad)
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
        raise AssertionError(f"{what}: {bad[0].size} entries differ; first at {i}: "
                             f"{a[i]!r} vs {b[i]!r}; max rel diff {rel:.3e}")


def compare_engines(ga, oa, sysm, label, conn=True):
    for p in range(len(sysm.particles)):
        sg, so = ga.state(p), oa.state(p)
        for k in FIELDS:
            assert_bitwise(sg[k], so[k], f"{label} particle {p} {k}")
        if conn:
            cg, co = ga.connectivity(p), oa.connectivity(p)
            assert cg.shape == co.shape and np.array_equal(cg, co), \
                f"{label} particle {p} connectivity: {cg.shape} vs {co.shape}"
    lg, rg = ga.wall()
    lo, ro = oa.wall()
    assert_bitwise(lg, lo, f"{label} wall planes")
    cg, co = ga.counters(), oa.counters()
    for k in ("steps", "bond_updates", "pair_computations", "alive_bonds", "total_bonds"):
        assert cg[k] == co[k], f"{label} counter {k}: {cg[k]} vs {co[k]}"
    return rg, ro


def small_impact(seed=5, R_h=4.0, v=150.0, self_contact=1, walls=True, b_movable=1,
                 fracture=1, gap_h=1.0, reaction=1):
    """Two small spheres colliding fast inside a moving walled box: exercises
    bond breaking, self-contact, pair contact, wall contact and reaction."""
    h = 1e-4
    s = gen.two_sphere_impact(R=R_h * h, h=h, v=v, seed=seed, gap_h=gap_h, b_movable=b_movable,
                              self_contact=self_contact, enable_fracture=fracture)
    if walls:
        lo = -R_h * h - 0.5 * h
        hi = 2 * R_h * h + gap_h * h + R_h * h + 0.5 * h
        s.allow_wall = 1
        s.compute_wall_reaction = reaction
        s.wall_lrtb = (lo, -R_h * h - 0.6 * h, -R_h * h - 0.7 * h, hi, R_h * h + 0.6 * h, R_h * h + 0.7 * h)
        s.wall_speed = (0.0, 5.0, 0.0, -20.0, 0.0, -3.0)
    rng = np.random.default_rng(seed + 100)
    B = s.particles[1]
    B.vel = rng.uniform(-5, 5, size=(B.n, 3))
    return s


def mixed_system(seed=7):
    """Several shapes with mixed flags, clamped nodes, gradient extforce, walls."""
    rng = np.random.default_rng(seed)
    h = 1e-4
    parts = []
    shapes = [gen.lattice_sphere(3.2), gen.lattice_cube(2.0), gen.lattice_convex(3.6, rng),
              gen.lattice_sphere(2.6)]
    centres = [(0, 0, 0), (6.3e-4, 0.2e-4, 0), (0.1e-4, 7.0e-4, 0.3e-4), (6.2e-4, 6.9e-4, -0.2e-4)]
    flags = [(1, 1, 1), (1, 0, 1), (0, 1, 1), (1, 1, 0)]
    for L, c, (mv, br, st) in zip(shapes, centres, flags):
        p = gen.make_particle(L, h, np.array(c), rng, movable=mv, breakable=br, stoppable=st)
        p.vel = rng.uniform(-2, 2, size=(p.n, 3))
        p.acc = rng.uniform(-1e4, 1e4, size=(p.n, 3))
        p.disp = rng.uniform(-1e-6, 1e-6, size=(p.n, 3))
        p.ext = np.zeros((p.n, 3))
        p.ext[:, 2] = -9.81 * p.rho
        parts.append(p)
    parts[0].clamped = np.array([0, 5, 17], dtype=np.uint32)
    Rc = 0.9 * h
    # centres of mass drift toward each other
    parts[0].vel[:, 0] += 60
    parts[1].vel[:, 0] -= 60
    parts[2].vel[:, 1] -= 60
    parts[3].vel[:, 1] -= 30
    s = gen.System(particles=parts, dt=1e-8, contact_rad=Rc,
                   normal_stiffness=gen.normal_stiffness(72e9, 0.25, Rc), friction_coefficient=0.3,
                   damping_ratio=0.1, enable_fracture=1, self_contact=1, allow_wall=1,
                   compute_wall_reaction=1, gradient_extforce=1, extforce_maxstep=37,
                   wall_lrtb=(-4.2e-4, -4.1e-4, -3.6e-4, 11.3e-4, 11.4e-4, 3.7e-4),
                   wall_speed=(1.0, 0.0, 0.0, 0.0, -2.0, -1.0), name="mixed")
    return s


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


def candidate_schedule(system, nsteps, oracle_lib):
    """The steps at which a job rebuilds its contact candidate lists and the
    longest list each rebuild needs, from oracle states (CPU): a list holds
    the nodes j of the job within rb = R_c (1 + 1e-9) + 0.3 R_c of node i,
    same-particle ones only with self-contact, measured on the displacement
    after the step's position half; lists are rebuilt at step 1 and whenever
    a node has moved more than half the skin since the last build."""
    Rc = system.contact_rad
    skin = 0.3 * Rc
    rb = Rc * (1 + 1e-9) + skin + 1e-300
    half = 0.5 * skin * (1 - 1e-6)
    X = np.concatenate([p.pos for p in system.particles])
    pid = np.concatenate([np.full(p.n, i) for i, p in enumerate(system.particles)])
    selfc = bool(system.self_contact and system.enable_fracture)
    e = oracle_lib.engine(system)
    ub, out = None, []
    try:
        for k in range(1, nsteps + 1):
            e.step(1)
            U = np.concatenate([e.state(p)["disp"] for p in range(len(system.particles))])
            if ub is not None and ((U - ub) ** 2).sum(1).max() <= half * half:
                continue
            ub = U.copy()
            d = (U[None, :, :] - U[:, None, :]) + (X[None, :, :] - X[:, None, :])
            dn = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
            ok = dn <= rb
            np.fill_diagonal(ok, False)
            if not selfc:
                ok &= pid[None, :] != pid[:, None]
            out.append((k, int(ok.sum(1).max())))
    finally:
        e.close()
    return out


def passing_spheres(Rc_h=1.8, v=1500.0, seed=21):
    """Two spheres with a contact spring too soft to stop them (K_n 1e20):
    they pass through each other, so the nodes of one come within the list
    radius of the other's nodes in growing numbers."""
    h = 1e-4
    s = gen.two_sphere_impact(R=4 * h, h=h, v=v, seed=seed, gap_h=1.0, b_movable=1, self_contact=0)
    s.contact_rad = Rc_h * h
    s.normal_stiffness = 1e20
    return s


def wheel_on_bulk(seed=31, wheel_y_h=9.0, wheel_vy=0.0, walls=False, self_contact=0):
    """A small sphere (particle 0, the wheel) above a cube (particle 1, the
    bulk), and a taller block (particle 2) off to the side in x: the setting
    of the wheel reset (timeloop.h:496-535).  Every node carries a seeded
    displacement; velocities are zero unless wheel_vy is set."""
    h = 1e-4
    rng = np.random.default_rng(seed)
    W = gen.make_particle(gen.lattice_sphere(3.0), h, np.array([0.0, wheel_y_h * h, 0.0]), rng)
    B = gen.make_particle(gen.lattice_cube(3.0), h, np.array([0.0, 0.0, 0.0]), rng)
    T = gen.make_particle(gen.lattice_cube(2.0), h, np.array([12 * h, 2 * h, 0.0]), rng)
    for p in (W, B, T):
        p.disp = rng.uniform(-2e-8, 2e-8, size=(p.n, 3))
        p.vel = np.zeros((p.n, 3))
    W.vel[:, 1] = wheel_vy
    Rc = 0.9 * h
    kw = dict(dt=1e-8, contact_rad=Rc, normal_stiffness=gen.contact_stiffness(Rc),
              friction_coefficient=0.3, damping_ratio=0.1, enable_fracture=1, self_contact=self_contact,
              allow_damping=1, allow_friction=1, name="wheel_on_bulk")
    if walls:
        kw.update(allow_wall=1, compute_wall_reaction=1,
                  wall_lrtb=(-5 * h, -4 * h, -4 * h, 15 * h, 14 * h, 4 * h),
                  wall_speed=(0.0, 0.0, 0.0, 0.0, -2.0, 0.0))
    return gen.System(particles=[W, B, T], **kw)


def unbonded(pos, like, vel=None):
    """A particle of the given nodes with no bonds (empty Connectivity), the
    material of `like`."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    h3 = float(like.vol[0])
    p = gen.Particle(pos=pos, vol=np.full(pos.shape[0], h3), conn=np.zeros((0, 2), dtype=np.uint32),
                     delta=like.delta, rho=like.rho, cnot=like.cnot, snot=like.snot)
    p.vel = np.zeros_like(pos) if vel is None else np.asarray(vel, dtype=np.float64).reshape(-1, 3)
    return p


def degenerate_particles(seed=41, walls=True):
    """Particles at the edges of the input space, all in contact within a few
    hundred steps: a bonded sphere (0); a single node flying into it (1);
    three unbonded nodes 2h apart catching it from behind (2); a 27-node cube
    with an empty Connectivity whose nodes collide with each other and with
    the top wall (3, self-contact among unbonded nodes); a bonded sphere one
    of whose nodes sits 1.6h off its surface with its bonds removed (4: a
    row of length zero among full ones)."""
    h = 1e-4
    rng = np.random.default_rng(seed)
    A = gen.make_particle(gen.lattice_sphere(4.0), h, np.array([0.0, 0.0, 0.0]), rng)
    A.vel = np.zeros((A.n, 3))
    A.vel[:, 0] = 20.0
    B = unbonded([[5.2 * h, 0.2 * h, 0.1 * h]], A, vel=[[-10.0, 0.0, 0.0]])
    C = unbonded([[-5.2 * h, 0.0, 0.0], [-5.2 * h, 2 * h, 0.0], [-5.2 * h, 0.0, 2 * h]], A,
                 vel=[[30.0, 0.0, 0.0]] * 3)
    L = gen.lattice_cube(1.0)
    D = unbonded(np.array([0.0, 9.0 * h, 0.0]) + h * (L + rng.uniform(-0.05, 0.05, size=L.shape)), A,
                 vel=rng.uniform(-6.0, 6.0, size=(L.shape[0], 3)))
    E = gen.make_particle(gen.lattice_sphere(3.0), h, np.array([0.0, -9.0 * h, 0.0]), rng)
    top = int(np.argmax(E.pos[:, 2]))
    E.pos[top, 2] += 1.6 * h
    E.conn = E.conn[(E.conn[:, 0] != top) & (E.conn[:, 1] != top)]
    E.vel = np.zeros((E.n, 3))
    E.vel[top, 2] = -40.0
    Rc = 0.9 * h
    kw = dict(dt=1e-8, contact_rad=Rc, normal_stiffness=gen.contact_stiffness(Rc),
              friction_coefficient=0.3, damping_ratio=0.1, enable_fracture=1, self_contact=1,
              allow_damping=1, allow_friction=1, name="degenerate_particles")
    if walls:
        kw.update(allow_wall=1, compute_wall_reaction=1,
                  wall_lrtb=(-12 * h, -14 * h, -6 * h, 12 * h, 10.6 * h, 6 * h),
                  wall_speed=(0.0, 0.0, 0.0, 0.0, -1.0, 0.0))
    return gen.System(particles=[A, B, C, D, E], **kw)
