"""The bond kernel's fast-path sqrt/division against the IEEE operators.

The library evaluates sqrt and division with the backend's own instruction
sequences minus their operand-rescaling wrappers, inside a guarded domain
(operands 0 or of magnitude in [2^-600, 2^600]).  These tests check, on the
GPU, bit for bit: the sequences against IEEE `/` and `sqrt` on random
operands spanning the whole domain, and that a run whose displacements leave
the domain switches to the IEEE operators and still matches the oracle.
"""
import numpy as np
import pytest

from helpers import compare_engines
from peridem_amd import abi, gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fast_division_and_sqrt_match_ieee(seed):
    n = 1 << 27
    print(f"[math] selftest n={n} seed={seed}", flush=True)
    md, ms = abi.selftest_math(n, seed)
    assert md == 0 and ms == 0, (md, ms)


def test_out_of_domain_displacement_switches_to_ieee(oracle_lib):
    s = gen.two_sphere_impact(R=4e-4, h=1e-4, seed=4, b_movable=1)
    s.particles[1].disp = np.zeros((s.particles[1].n, 3))
    s.particles[1].disp[7, 1] = 1e-200          # below the fast domain
    g = abi.gpu_engine(s)
    o = oracle_lib.engine(s)
    try:
        assert not g.math_mode_safe()
        g.step(3)
        o.step(3)
        assert g.math_mode_safe()
        compare_engines(g, o, s, "out-of-domain displacement")
    finally:
        g.close()
        o.close()
