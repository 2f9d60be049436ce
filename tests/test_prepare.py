"""Host-side job layout of the C-ABI library (pd_job_prepare and its footprint),
which runs without a GPU: argument checks return the library's error codes
with a message instead of ending the process (the reference prints and
continues or calls exit, src/debug/err.h:21-24), and footprints follow the
256-node block layout."""
import numpy as np
import pytest

from helpers import small_jobs
from peridem_amd import abi, gen


def one_sphere(seed=3, **kw):
    rng = np.random.default_rng(seed)
    p = gen.make_particle(gen.lattice_sphere(4.0), 1e-4, np.zeros(3), rng, **kw)
    return gen.System(particles=[p], dt=1e-8)


def test_footprint_is_whole_blocks_and_covers_every_bond():
    for s in small_jobs(4, seed=5) + [one_sphere()]:
        b = abi.prepare_job(s)
        try:
            slots, entries, nbytes = b.footprint()
        finally:
            b.free()
        assert slots % 256 == 0 and s.nodes <= slots < s.nodes + 256 * len(s.particles)
        assert entries % 64 == 0 and entries >= s.bonds
        assert nbytes > 6 * entries


def test_connectivity_entry_out_of_range_is_an_error():
    s = one_sphere()
    s.particles[0].conn = s.particles[0].conn.copy()
    s.particles[0].conn[7, 1] = s.particles[0].n      # one past the last node
    with pytest.raises(abi.PDError, match="Connectivity entry 15 out of range"):
        abi.prepare_job(s)


def test_particle_without_nodes_is_an_error():
    s = one_sphere()
    p = s.particles[0]
    p.pos = p.pos[:0]
    p.vol = p.vol[:0]
    p.conn = p.conn[:0]
    with pytest.raises(abi.PDError, match="needs nnodes"):
        abi.prepare_job(s)


def test_torque_axis_out_of_range_is_an_error():
    s = one_sphere()
    s.particles[0].torque_axis = 3
    with pytest.raises(abi.PDError, match="torque_axis"):
        abi.prepare_job(s)


def test_boundary_node_listed_twice_is_unsupported():
    s = one_sphere()
    s.nl_bdry_only = 1
    s.particles[0].bdry = np.array([0, 5, 9, 5], dtype=np.uint32)
    with pytest.raises(abi.PDError, match="PD_E_UNSUPPORTED"):
        abi.prepare_job(s)


def test_wide_horizon_still_prepares():
    """A horizon of 5h makes 256-node blocks reference more nodes than the
    staged bond kernel's table holds; the job still lays out (the pool then
    runs the unstaged kernel on it)."""
    s = one_sphere(horizon=5.0)
    b = abi.prepare_job(s)
    try:
        slots, entries, _ = b.footprint()
    finally:
        b.free()
    assert slots >= s.nodes and entries >= s.bonds
