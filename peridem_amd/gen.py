"""Synthetic setups for the explicit step: seeded jittered-lattice particles.

The reference builds its setup file offline (exp_dict.py: gmsh meshes,
`cKDTree.query_pairs(delta)` bonds sorted lexicographically, exp_dict.py:544-583,
written by Experiment.save, exp_dict.py:2119-2186).  This module produces
the same data model without gmsh: nodes are the points of a cubic lattice
(spacing h) inside a shape, jittered uniformly by +-0.05h with a seeded
generator, each with volume h^3; bonds are all pairs with |X_j - X_i| <= delta,
emitted as (i<j) rows in lexicographic order, so conn2NArr rows come out
ascending exactly as they do for the reference's sorted query_pairs output.

Material constants follow the 3-D PMB closed forms of material_dict.py:266-277
(peridem_3d): c0 = 18K/(pi delta^4), s0 = sqrt(5 Gc / (9 K delta)), with the
bulk modulus K taken from the particle's own E and nu.  The contact normal
stiffness is K_n = 18K/(pi R_c^5), i.e. c0 evaluated at delta = R_c divided by
R_c (contact.h:775-776), with K = 2e9 Pa as the reference's aggregate example
takes it (see contact_stiffness).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from .abi import ParticleDesc, RunDesc


# ---------------------------------------------------------------------------
# material and contact constants
def bulk_modulus(E, nu):
    return E / (3.0 * (1.0 - 2.0 * nu))


def pmb3d(E, nu, Gc, delta):
    """(cnot, snot) of the 3-D bond-based prototype microelastic brittle model."""
    K = bulk_modulus(E, nu)
    return 18.0 * K / (math.pi * delta ** 4), math.sqrt(5.0 * Gc / (9.0 * K * delta))


def normal_stiffness(E, nu, Rc):
    """18 K / (pi R_c^5) with the particle's own bulk modulus (c0 at delta = R_c,
    over R_c, contact.h:775-776)."""
    return 18.0 * bulk_modulus(E, nu) / (math.pi * Rc ** 5)


# The reference's aggregate example takes the contact stiffness from the
# peridem_3d material whatever the particles are made of:
#   normal_stiffness = material_dict.peridem_3d(contact_radius).cnot / contact_radius
# (examples/3D-HPC-aggregate-particle-breakage/setup.py:160; peridem_3d has a
# bulk modulus of 2e9 Pa, material_dict.py:266-277).  The workloads below use
# that convention: with the particles' own 48e9 Pa the spring is 24x stiffer,
# and two spheres meeting at 10 m/s blow up within a few hundred steps of
# dt = 1e-8 s (in the reference algorithm as in this library), while with
# 2e9 Pa they bounce (tests/test_generator.py).
CONTACT_BULK_MODULUS = 2e9


def contact_stiffness(Rc, K=CONTACT_BULK_MODULUS):
    """peridem_3d(R_c).cnot / R_c = 18 K / (pi R_c^5), K = 2e9 Pa."""
    return 18.0 * K / (math.pi * Rc ** 5)


# ---------------------------------------------------------------------------
@dataclass
class Particle:
    pos: np.ndarray
    vol: np.ndarray
    conn: np.ndarray
    delta: float
    rho: float
    cnot: float
    snot: float
    disp: np.ndarray | None = None
    vel: np.ndarray | None = None
    acc: np.ndarray | None = None
    ext: np.ndarray | None = None
    clamped: np.ndarray | None = None
    bdry: np.ndarray | None = None
    movable: int = 1
    breakable: int = 1
    stoppable: int = 1
    torque_axis: int = 2
    torque_val: float = 0.0
    instance: int = 0

    @property
    def n(self):
        return int(self.pos.shape[0])

    @property
    def nbonds(self):
        return 2 * int(self.conn.shape[0])

    def desc(self, keep):
        def f64(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        def u32(a):
            if a is None or len(a) == 0:
                return None, 0
            a = np.ascontiguousarray(a, dtype=np.uint32)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_uint32)), a.size

        d = ParticleDesc()
        d.nnodes = self.n
        d.instance = self.instance
        d.pos, d.vol = f64(self.pos), f64(self.vol)
        d.disp, d.vel, d.acc, d.extforce = f64(self.disp), f64(self.vel), f64(self.acc), f64(self.ext)
        cp, cn = u32(self.conn.reshape(-1))
        d.conn, d.nconn = cp, cn // 2
        d.clamped, d.nclamped = u32(self.clamped)
        d.bdry, d.nbdry = u32(self.bdry)
        d.delta, d.rho, d.cnot, d.snot = self.delta, self.rho, self.cnot, self.snot
        d.movable, d.breakable, d.stoppable = self.movable, self.breakable, self.stoppable
        d.torque_axis, d.torque_val = self.torque_axis, self.torque_val
        return d


@dataclass
class System:
    particles: list
    dt: float = 1e-8
    contact_rad: float = 0.0
    normal_stiffness: float = 0.0
    friction_coefficient: float = 0.3
    damping_ratio: float = 0.1
    enable_fracture: int = 1
    self_contact: int = 0
    allow_damping: int = 1
    allow_friction: int = 1
    nl_bdry_only: int = 0
    compute_wall_reaction: int = 0
    gradient_extforce: int = 0
    extforce_maxstep: int = 0
    enable_torque: int = 0
    allow_wall: int = 0
    wall_lrtb: tuple = (0, 0, 0, 0, 0, 0)
    wall_speed: tuple = (0, 0, 0, 0, 0, 0)
    first_step: int = 1
    name: str = ""
    meta: dict = field(default_factory=dict)

    def run_desc(self):
        r = RunDesc()
        for k in ("dt", "contact_rad", "normal_stiffness", "friction_coefficient", "damping_ratio"):
            setattr(r, k, float(getattr(self, k)))
        for k in ("enable_fracture", "self_contact", "allow_damping", "allow_friction",
                  "nl_bdry_only", "compute_wall_reaction", "gradient_extforce",
                  "extforce_maxstep", "enable_torque", "allow_wall", "first_step"):
            setattr(r, k, int(getattr(self, k)))
        for k in range(6):
            r.wall_lrtb[k] = float(self.wall_lrtb[k])
            r.wall_speed[k] = float(self.wall_speed[k])
        return r

    @property
    def nodes(self):
        return sum(p.n for p in self.particles)

    @property
    def bonds(self):
        return sum(p.nbonds for p in self.particles)

    def movable_bonds(self):
        return sum(p.nbonds for p in self.particles if p.movable)


# ---------------------------------------------------------------------------
# lattice shapes (integer coordinates, C order)
def lattice_sphere(r_over_h):
    m = int(math.floor(r_over_h))
    ax = np.arange(-m, m + 1)
    I, J, K = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    keep = (pts ** 2).sum(1) <= r_over_h ** 2 + 1e-9
    return pts[keep]


def lattice_cube(half_over_h):
    m = int(math.floor(half_over_h))
    ax = np.arange(-m, m + 1)
    I, J, K = np.meshgrid(ax, ax, ax, indexing="ij")
    return np.stack([I.ravel(), J.ravel(), K.ravel()], 1)


def lattice_convex(r_over_h, rng, nplanes=14):
    """Random convex polyhedron: a ball cut by random tangent-ish planes."""
    pts = lattice_sphere(r_over_h)
    normals = rng.normal(size=(nplanes, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    offs = r_over_h * rng.uniform(0.55, 0.95, size=nplanes)
    keep = np.all(pts @ normals.T <= offs[None, :] + 1e-9, axis=1)
    return pts[keep]


def bonds_lattice(L, X, delta, h, jitter=0.05):
    """All pairs (i<j) with |X_j-X_i| <= delta, lexicographically sorted.

    L: integer lattice coordinates of the nodes, X: jittered positions.
    Candidates are lattice offsets up to delta/h plus the largest possible
    jitter change of a distance; the exact test is on the jittered positions.
    """
    n = L.shape[0]
    lo = L.min(0)
    Ls = L - lo
    reach = int(math.ceil(delta / h + 2 * jitter * math.sqrt(3) + 1e-9))
    pad = reach + 1
    shape = tuple(int(v) for v in (Ls.max(0) + 1 + 2 * pad))
    grid = -np.ones(shape, dtype=np.int64)
    Lp = Ls + pad
    grid[Lp[:, 0], Lp[:, 1], Lp[:, 2]] = np.arange(n)
    lim2 = (delta / h + 2 * jitter * math.sqrt(3)) ** 2 + 1e-9
    keys = []
    rng_ = range(-reach, reach + 1)
    for ox in rng_:
        for oy in rng_:
            for oz in rng_:
                o = (ox, oy, oz)
                if o <= (0, 0, 0):
                    continue
                if ox * ox + oy * oy + oz * oz > lim2:
                    continue
                j = grid[Lp[:, 0] + ox, Lp[:, 1] + oy, Lp[:, 2] + oz]
                ok = j >= 0
                i = np.nonzero(ok)[0]
                j = j[ok]
                d = X[j] - X[i]
                dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
                sel = dist <= delta
                i, j = i[sel], j[sel]
                a, b = np.minimum(i, j), np.maximum(i, j)
                keys.append((a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64))
    if not keys:
        return np.zeros((0, 2), dtype=np.uint32)
    k = np.sort(np.concatenate(keys))
    return np.stack([(k >> np.uint64(32)).astype(np.uint32), (k & np.uint64(0xFFFFFFFF)).astype(np.uint32)], 1)


def make_particle(L, h, center, rng, E=72e9, nu=0.25, rho=2440.0, Gc=135.0, horizon=3.0,
                  jitter=0.05, **kw):
    X = center[None, :] + h * (L + rng.uniform(-jitter, jitter, size=L.shape))
    delta = horizon * h
    cnot, snot = pmb3d(E, nu, Gc, delta)
    conn = bonds_lattice(L, X, delta, h, jitter)
    n = L.shape[0]
    return Particle(pos=X, vol=np.full(n, h ** 3), conn=conn, delta=delta, rho=rho, cnot=cnot,
                    snot=snot, **kw)


# ---------------------------------------------------------------------------
# the named workloads
def two_sphere_impact(R=1e-3, h=1e-4, v=20.0, seed=1, E=72e9, nu=0.25, rho=2440.0, Gc=135.0,
                      gap_h=1.1, b_movable=0, **sys_kw):
    """Sphere A at v m/s toward a fixed sphere B (surface gap gap_h*h)."""
    rng = np.random.default_rng(seed)
    L = lattice_sphere(R / h)
    A = make_particle(L, h, np.array([0.0, 0.0, 0.0]), rng, E, nu, rho, Gc)
    B = make_particle(L, h, np.array([2 * R + gap_h * h, 0.0, 0.0]), rng, E, nu, rho, Gc,
                      movable=b_movable)
    A.vel = np.zeros((A.n, 3))
    A.vel[:, 0] = v
    Rc = 0.9 * h
    kw = dict(dt=1e-8, contact_rad=Rc, normal_stiffness=contact_stiffness(Rc),
              friction_coefficient=0.3, damping_ratio=0.1, enable_fracture=1, self_contact=0,
              name="two_sphere_impact")
    kw.update(sys_kw)
    return System(particles=[A, B], **kw)


def _jittered_grid_centres(n, R, box, gap, rng):
    """n centres on an m^3 grid filling the box, each moved by a seeded uniform
    jitter that keeps every pair >= 2R + gap apart and every sphere >= gap
    from the walls.  (Random sequential addition jams near 38% packing, below
    what a 30%-packed box with contact margins needs.)"""
    m = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
    pitch = box / m
    jit = 0.5 * (pitch - 2 * R - gap)
    wall_jit = 0.5 * pitch - R - gap
    j = min(jit, wall_jit)
    if j < 0:
        raise RuntimeError(f"{n} spheres of radius {R} do not fit a {box} box")
    cells = [(i, k, l) for l in range(m) for k in range(m) for i in range(m)][:n]
    cs = np.array([[(i + 0.5) * pitch, (k + 0.5) * pitch, (l + 0.5) * pitch] for i, k, l in cells])
    return cs + rng.uniform(-j, j, size=cs.shape)


def gravity_settling(nspheres=125, R=1e-3, nodes_per_sphere=8000, box=1.2e-2, seed=2,
                     E=72e9, nu=0.25, rho=2440.0, Gc=135.0, g=9.81, sphere_speed=0.0,
                     node_speed=0.0, **sys_kw):
    """Spheres at seeded random non-overlapping centres in a walled box under gravity.

    With sphere_speed = node_speed = 0 every node of a sphere keeps the same
    displacement, so every bond keeps exactly zero stretch.  sphere_speed adds
    a seeded rigid velocity per sphere (uniform in +-sphere_speed per axis)
    and node_speed a seeded per-node velocity (uniform in +-node_speed), which
    makes every bond carry stretch.
    """
    rng = np.random.default_rng(seed)
    h = R * (4.0 * math.pi / 3.0 / nodes_per_sphere) ** (1.0 / 3.0)
    L = lattice_sphere(R / h)
    Rc = 0.9 * h
    cs = _jittered_grid_centres(nspheres, R, box, 2 * h, rng)
    parts = []
    vrng = np.random.default_rng(seed + 7919)
    for c in cs:
        p = make_particle(L, h, c, rng, E, nu, rho, Gc)
        p.ext = np.zeros((p.n, 3))
        p.ext[:, 2] = -g * rho
        if sphere_speed or node_speed:
            p.vel = (vrng.uniform(-sphere_speed, sphere_speed, size=(1, 3)) +
                     vrng.uniform(-node_speed, node_speed, size=(p.n, 3)))
        parts.append(p)
    kw = dict(dt=1e-8, contact_rad=Rc, normal_stiffness=contact_stiffness(Rc),
              friction_coefficient=0.3, damping_ratio=0.1, enable_fracture=1, self_contact=0,
              allow_wall=1, wall_lrtb=(0, 0, 0, box, box, box), name="gravity_settling")
    kw.update(sys_kw)
    s = System(particles=parts, **kw)
    s.meta["h"] = h
    return s


def _rsa_centres(radii, frac, rng, margin, gap):
    """Random sequential placement, largest first, of spheres of the given
    radii in a cube whose side puts their volume at `frac` of it: each centre
    is drawn uniformly at least radius + margin from the walls and kept if it
    is at least r_i + r_j + gap from every placed one (the first free one of
    batches of draws); the cube grows by 2% and placement restarts when a
    sphere finds no room in 20000 draws."""
    radii = np.asarray(radii, dtype=float)
    order = np.argsort(-radii, kind="stable")
    side = (4.0 / 3.0 * math.pi * np.sum(radii ** 3) / frac) ** (1.0 / 3.0)
    while True:
        placed = np.zeros((0, 3))
        prad = np.zeros(0)
        ok = True
        for k in order:
            r = radii[k]
            lo, hi = r + margin, side - r - margin
            for _ in range(40):   # 40 batches of 500 draws, the first free one kept
                cand = rng.uniform(lo, hi, size=(500, 3))
                if placed.shape[0]:
                    dist = np.linalg.norm(cand[:, None, :] - placed[None, :, :], axis=2)
                    free = np.all(dist >= prad[None, :] + r + gap, axis=1)
                else:
                    free = np.ones(len(cand), dtype=bool)
                if free.any():
                    placed = np.vstack([placed, cand[np.argmax(free)]])
                    prad = np.append(prad, r)
                    break
            else:
                ok = False
                break
        if ok:
            out = np.zeros((len(radii), 3))
            out[order] = placed
            return out, side
        side *= 1.02


def crushing_aggregate(nparticles=512, nodes_range=(2000, 12000), seed=3, E=72e9, nu=0.25,
                       rho=2440.0, Gc=135.0, h=1e-4, top_speed=-5.0, packing=0.35, **sys_kw):
    """Mixed cubes, spheres and convex polyhedra of node counts drawn uniformly
    from nodes_range (512 of 2k-12k nodes: ~3.6M nodes) placed at random without
    overlap (their bounding spheres fill `packing` of the box), top wall
    moving down.

    h = 1e-4 m as in the other workloads: the wall contact spring (contact.h:385-502,
    K_n = 18K/(pi R_c^5), R_c = 0.9h) is stiffer per unit volume as R_c shrinks,
    and at h = 6e-5 m dt = 1e-8 s is beyond its stability limit -- a particle
    touching the moving wall blows up within a hundred steps, in the reference
    algorithm as much as here."""
    rng = np.random.default_rng(seed)
    parts = []
    sizes = []
    for k in range(nparticles):
        target = float(rng.uniform(nodes_range[0], nodes_range[1]))
        kind = rng.integers(0, 3)
        if kind == 0:
            half = 0.5 * target ** (1 / 3) - 0.5
            L = lattice_cube(half)
        elif kind == 1:
            L = lattice_sphere((3 * target / (4 * math.pi)) ** (1 / 3))
        else:
            L = lattice_convex((3 * target / (4 * math.pi)) ** (1 / 3) * 1.25, rng)
        sizes.append(L)
    # bounding radius of each lattice (in h) plus the jitter
    radii = np.array([float(np.sqrt((np.asarray(L, float) ** 2).sum(1)).max()) + 0.1 for L in sizes]) * h
    Rc = 0.9 * h
    centres, side = _rsa_centres(radii, packing, rng, margin=2 * h, gap=1.5 * Rc)
    for L, c in zip(sizes, centres):
        parts.append(make_particle(L, h, c, rng, E, nu, rho, Gc))
    kw = dict(dt=1e-8, contact_rad=Rc, normal_stiffness=contact_stiffness(Rc),
              friction_coefficient=0.3, damping_ratio=0.1, enable_fracture=1, self_contact=1,
              allow_wall=1, compute_wall_reaction=1,
              wall_lrtb=(0, 0, 0, side, side, side),
              wall_speed=(0, 0, 0, 0, 0, top_speed), name="crushing_aggregate")
    kw.update(sys_kw)
    return System(particles=parts, **kw)


def impact_pair(nA, nB, v, E, seed=1, h=1e-4, nu=0.25, rho=2440.0, Gc=135.0, gap_h=1.1,
                **sys_kw):
    """Two free spheres of about nA and nB lattice nodes (spacing h); A moves at
    v m/s along +x toward B.  Material constants, contact stiffness and s0
    follow from E (the per-instance parameters of a batch)."""
    rng = np.random.default_rng(seed)
    rA = (3.0 * nA / (4.0 * math.pi)) ** (1.0 / 3.0)
    rB = (3.0 * nB / (4.0 * math.pi)) ** (1.0 / 3.0)
    LA, LB = lattice_sphere(rA), lattice_sphere(rB)
    A = make_particle(LA, h, np.array([0.0, 0.0, 0.0]), rng, E, nu, rho, Gc)
    B = make_particle(LB, h, np.array([(rA + rB + gap_h) * h, 0.0, 0.0]), rng, E, nu, rho, Gc)
    A.vel = np.zeros((A.n, 3))
    A.vel[:, 0] = v
    Rc = 0.9 * h
    kw = dict(dt=1e-8, contact_rad=Rc, normal_stiffness=contact_stiffness(Rc),
              friction_coefficient=0.3, damping_ratio=0.1, enable_fracture=1, self_contact=1,
              name="impact_pair")
    kw.update(sys_kw)
    s = System(particles=[A, B], **kw)
    s.meta.update(nA=A.n, nB=B.n, v=v, E=E, seed=seed)
    return s


def impact_batch(n=256, seed=4, nodes_range=(500, 20000), speed_range=(5.0, 50.0),
                 E_range=(20e9, 100e9), **kw):
    """n independent two-sphere impacts: node counts per particle log-uniform
    on nodes_range, impact speed and Young's modulus uniform per instance."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(nodes_range[0]), np.log(nodes_range[1])
    out = []
    for k in range(n):
        nA, nB = np.exp(rng.uniform(lo, hi, size=2))
        v = float(rng.uniform(*speed_range))
        E = float(rng.uniform(*E_range))
        out.append(impact_pair(float(nA), float(nB), v, E, seed=int(seed * 100003 + k), **kw))
    return out


# ---------------------------------------------------------------------------
# served aggregate-breakage jobs: particles are translated copies of a few
# cached lattice templates, so a job with 64 particles builds in milliseconds
_TEMPLATES = {}


def _template(kind, nodes, h, seed, E=72e9, nu=0.25, rho=2440.0, Gc=135.0):
    key = (kind, int(nodes), h, seed, E)
    if key not in _TEMPLATES:
        rng = np.random.default_rng(seed)
        if kind == "cube":
            L = lattice_cube(0.5 * nodes ** (1.0 / 3.0) - 0.5)
        else:
            L = lattice_sphere((3.0 * nodes / (4.0 * math.pi)) ** (1.0 / 3.0))
        p = make_particle(L, h, np.zeros(3), rng, E, nu, rho, Gc)
        _TEMPLATES[key] = p
    return _TEMPLATES[key]


def aggregate_job(job_id, nparticles_range=(2, 64), nodes_range=(2000, 8000),
                  steps_range=(5000, 50000), h=1e-4, speed=8.0, buckets=6, seed=7, **sys_kw):
    """One served job: 2-64 seeded spheres and cubes of 2k-8k nodes on a
    jittered grid inside walls, each thrown toward the cluster centre, with a
    seeded step budget.  Node counts are drawn log-uniform and rounded to one
    of `buckets` sizes per shape so the lattices can be reused."""
    rng = np.random.default_rng(seed * 1000003 + job_id)
    n = int(rng.integers(nparticles_range[0], nparticles_range[1] + 1))
    lo, hi = np.log(nodes_range[0]), np.log(nodes_range[1])
    sizes = np.exp(np.linspace(lo, hi, buckets))
    picks = []
    for _ in range(n):
        kind = "cube" if rng.random() < 0.5 else "sphere"
        b = int(np.argmin(np.abs(sizes - np.exp(rng.uniform(lo, hi)))))
        picks.append(_template(kind, sizes[b], h, seed=b * 2 + (kind == "cube")))
    ext = max(float(np.abs(t.pos).max()) for t in picks)
    pitch = 2 * ext + 2.5 * h
    m = int(math.ceil(n ** (1.0 / 3.0)))
    parts = []
    cells = [(i, j, k) for k in range(m) for j in range(m) for i in range(m)][:n]
    centre = np.full(3, 0.5 * m * pitch)
    for t, (i, j, k) in zip(picks, cells):
        c = np.array([(i + 0.5) * pitch, (j + 0.5) * pitch, (k + 0.5) * pitch])
        c = c + rng.uniform(-0.2, 0.2, size=3) * h
        p = Particle(pos=t.pos + c, vol=t.vol, conn=t.conn, delta=t.delta, rho=t.rho, cnot=t.cnot,
                     snot=t.snot)
        d = centre - c
        nd = float(np.linalg.norm(d))
        v = (d / nd if nd > 0 else rng.normal(size=3)) * speed * rng.uniform(0.5, 1.0)
        p.vel = np.broadcast_to(v, (p.n, 3)).copy()
        parts.append(p)
    Rc = 0.9 * h
    box = m * pitch
    kw = dict(dt=1e-8, contact_rad=Rc, normal_stiffness=contact_stiffness(Rc),
              friction_coefficient=0.3, damping_ratio=0.1, enable_fracture=1, self_contact=1,
              allow_wall=1, wall_lrtb=(0, 0, 0, box, box, box), name="aggregate_job")
    kw.update(sys_kw)
    s = System(particles=parts, **kw)
    s.meta.update(job_id=job_id, steps=int(rng.integers(steps_range[0], steps_range[1] + 1)),
                  nparticles=n)
    return s
