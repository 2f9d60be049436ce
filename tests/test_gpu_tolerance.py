"""Tolerance arithmetic of the bond kernel (pd_set_arithmetic).

PD_ARITH_TOLERANCE takes |e| and |xi| from the hardware reciprocal square
root refined by one Newton-Raphson step.  The bar (the stated tolerance):
per step, from the same state, each node's force density is within
1e-12 x c0 sum_j V_j of the oracle's -- the force a stretch error of 1e-12
on every bond of its row would make (particle2.h:139-153: term c0 s V_j e/|e|)
-- and the broken-bond set is the oracle's for every bond whose stretch lies
farther than 1e-12 from s0.  Each step is teacher-forced: the GPU takes the
oracle's state (disp, vel, acc) before stepping, so the per-step error is
measured, not its accumulation.  Velocity and acceleration follow from the
force (a = f / rho, v += (a_old + a) dt / 2), so they carry the same bound
scaled; displacement and the position half are bitwise.
"""
import numpy as np
import pytest

from helpers import assert_bitwise, mixed_system, small_impact
from peridem_amd import abi

pytestmark = pytest.mark.gpu

TOL = 1e-12


def stiffness_scale(p):
    """c0 * sum of V_j over each node's initial row."""
    s = np.zeros(p.n)
    np.add.at(s, p.conn[:, 0], p.vol[p.conn[:, 1]])
    np.add.at(s, p.conn[:, 1], p.vol[p.conn[:, 0]])
    return p.cnot * s


@pytest.mark.parametrize("case", ["impact", "mixed"])
def test_tolerance_mode_per_step_against_the_oracle(oracle_lib, case):
    s = small_impact(seed=14) if case == "impact" else mixed_system(seed=5)
    g, o = abi.gpu_engine(s), oracle_lib.engine(s)
    g.set_arithmetic(abi.PD_ARITH_TOLERANCE)
    scale = [stiffness_scale(p) for p in s.particles]
    worst = 0.0
    try:
        for step in range(60):
            for p in range(len(s.particles)):
                so = o.state(p)
                g.set_state(p, disp=so["disp"], vel=so["vel"], acc=so["acc"])
            g.step(1)
            o.step(1)
            for p, P in enumerate(s.particles):
                sg, so = g.state(p), o.state(p)
                assert_bitwise(sg["disp"], so["disp"], f"step {step} particle {p} disp")
                lim = TOL * scale[p][:, None] + 1e-300
                df = np.abs(sg["force"] - so["force"])
                worst = max(worst, float((df / lim).max()))
                assert np.all(df <= lim), f"step {step} particle {p}: force off by {(df / lim).max():.3g} x the bound"
                if P.movable:
                    assert np.all(np.abs(sg["acc"] - so["acc"]) <= lim / P.rho)
                    assert np.all(np.abs(sg["vel"] - so["vel"]) <= lim / P.rho * s.dt)
                assert np.array_equal(g.connectivity(p), o.connectivity(p)), f"step {step} particle {p} bonds"
        assert o.counters()["alive_bonds"] < o.counters()["total_bonds"] or case == "mixed"
        print(f"[tolerance] {case}: worst force error {worst:.3g} x the bound over 60 steps", flush=True)
    finally:
        g.close()
        o.close()


def test_back_to_bitwise_arithmetic_is_bitwise_again(oracle_lib):
    """After teacher-forced steps in tolerance arithmetic, switching back to
    PD_ARITH_BITWISE and taking the oracle's state gives the oracle's bits."""
    s = small_impact(seed=15)
    g, o = abi.gpu_engine(s), oracle_lib.engine(s)
    g.set_arithmetic(abi.PD_ARITH_TOLERANCE)
    try:
        for phase, n in (("tolerance", 20), ("bitwise", 20)):
            if phase == "bitwise":
                g.set_arithmetic(abi.PD_ARITH_BITWISE)
            for _ in range(n):
                for p in range(len(s.particles)):
                    so = o.state(p)
                    g.set_state(p, disp=so["disp"], vel=so["vel"], acc=so["acc"])
                g.step(1)
                o.step(1)
        for p in range(len(s.particles)):
            sg, so = g.state(p), o.state(p)
            for k in ("disp", "vel", "acc", "force"):
                assert_bitwise(sg[k], so[k], f"bitwise again, particle {p} {k}")
            assert np.array_equal(g.connectivity(p), o.connectivity(p))
    finally:
        g.close()
        o.close()
