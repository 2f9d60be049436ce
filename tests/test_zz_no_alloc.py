"""The allocation-free math self-test against the allocating one.

pd_selftest_math_noalloc keeps its counters in the library's static device
array instead of allocating them; it must report what pd_selftest_math
reports for the same operands, and the array must start from zero on every
call (the second, smaller call would otherwise count the first call's
samples too).
"""
import pytest

from peridem_amd import abi

pytestmark = pytest.mark.gpu


def test_noalloc_selftest_matches_the_allocating_entry_point():
    for n, seed in ((1 << 22, 11), (123457, 12)):
        print(f"[noalloc] n={n} seed={seed}", flush=True)
        md, ms = abi.selftest_math(n, seed)
        nd, ns, ev = abi.selftest_math_noalloc(n, seed)
        assert (nd, ns) == (md, ms) == (0, 0)
        assert ev == n                          # every sample counted once, nothing left over


def test_noalloc_selftest_rejects_oversized_requests():
    with pytest.raises(abi.PDError):
        abi.selftest_math_noalloc((1 << 40) + 1, 1)
