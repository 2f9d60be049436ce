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
