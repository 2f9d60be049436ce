"""The synthetic setup generator: material constants, bond sets and the
configuration sizes the measurements quote.

The constants follow the 3-D PMB closed forms of material_dict.py:266-277:
K = E / (3 (1 - 2 nu)), c0 = 18 K / (pi delta^4), s0 = sqrt(5 Gc / (9 K delta)),
and the contact stiffness convention K_n = 18 K / (pi R_c^5) (contact.h:775-776).
"""
import math

import numpy as np

from peridem_amd import gen


def test_pmb_constants_of_the_two_sphere_configuration():
    # E = 72 GPa, nu = 1/4 -> K = 48 GPa; delta = 3h = 3e-4 m; Gc = 135 J/m^2
    c0, s0 = gen.pmb3d(72e9, 0.25, 135.0, 3e-4)
    K = 48e9
    assert gen.bulk_modulus(72e9, 0.25) == K
    assert math.isclose(c0, 18 * K / (math.pi * 3e-4 ** 4), rel_tol=1e-15)
    assert math.isclose(c0, 3.395e25, rel_tol=1e-3)
    assert math.isclose(s0, math.sqrt(5 * 135.0 / (9 * K * 3e-4)), rel_tol=1e-15)
    assert math.isclose(s0, 2.282e-3, rel_tol=1e-3)
    # K_n is c0 evaluated at delta = R_c, divided by R_c
    Rc = 9e-5
    c0_rc, _ = gen.pmb3d(72e9, 0.25, 135.0, Rc)
    assert math.isclose(gen.normal_stiffness(72e9, 0.25, Rc), c0_rc / Rc, rel_tol=1e-14)


def test_bonds_are_exactly_the_pairs_within_the_horizon():
    rng = np.random.default_rng(8)
    h = 1e-4
    p = gen.make_particle(gen.lattice_sphere(3.5), h, np.zeros(3), rng)
    X = p.pos
    d = np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    i, j = np.nonzero(np.triu(d <= p.delta, k=1))
    want = set(zip(i.tolist(), j.tolist()))
    got = set(map(tuple, p.conn.tolist()))
    assert got == want
    # emitted as (i < j) rows in lexicographic order, like the sorted query_pairs output
    assert np.all(p.conn[:, 0] < p.conn[:, 1])
    keys = p.conn[:, 0].astype(np.int64) * p.n + p.conn[:, 1]
    assert np.all(np.diff(keys) > 0)
    assert np.all(p.vol == h ** 3)


def test_headline_workload_sizes():
    """Seeded: the first two spheres of the 125-sphere workload (8,025 nodes
    each; the jitter decides the bonds at exactly |xi| = 3h)."""
    s = gen.gravity_settling(nspheres=2)
    assert [q.n for q in s.particles] == [8025, 8025]
    assert [q.nbonds for q in s.particles] == [741090, 740742]


def test_contact_stiffness_is_the_reference_examples():
    """normal_stiffness = peridem_3d(R_c).cnot / R_c
    (examples/3D-HPC-aggregate-particle-breakage/setup.py:160): bulk modulus
    2e9 Pa whatever the particles' material (material_dict.py:266-277)."""
    Rc = 9e-5
    assert math.isclose(gen.contact_stiffness(Rc), 18 * 2e9 / (math.pi * Rc ** 4) / Rc, rel_tol=1e-14)
    s = gen.two_sphere_impact(R=4e-4, h=1e-4, v=20.0, seed=1)
    assert s.normal_stiffness == gen.contact_stiffness(s.contact_rad)


def test_workload_contact_is_stable_at_the_workloads_time_step(oracle_lib):
    """Two spheres meeting at 10 m/s from just outside R_c, dt = 1e-8 s (the
    reference algorithm, in the oracle): with the workloads' contact stiffness
    they bounce -- speeds stay near 10 m/s and no bond breaks; with the 24x
    stiffer spring of the particles' own bulk modulus the same impact blows
    up (speeds beyond 1e3 m/s, bonds broken)."""
    out = {}
    for label, Kn in (("workload", None), ("particle K", gen.normal_stiffness(72e9, 0.25, 9e-5))):
        s = gen.two_sphere_impact(R=4e-4, h=1e-4, v=10.0, seed=3, gap_h=1.0, b_movable=1, self_contact=1)
        if Kn is not None:
            s.normal_stiffness = Kn
        o = oracle_lib.engine(s, prune=True)
        o.step(600)
        vmax = max(float(np.abs(o.state(p)["vel"]).max()) for p in range(2))
        c = o.counters()
        out[label] = (vmax, c["total_bonds"] - c["alive_bonds"], c["pair_computations"])
        o.close()
    v, broken, pairs = out["workload"]
    assert pairs > 0 and v < 20.0 and broken == 0, out
    v, broken, _ = out["particle K"]
    assert v > 1e3 and broken > 0, out
