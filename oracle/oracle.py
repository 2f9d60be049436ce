"""Python loader of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check the HIP library's results and by bench.py's CPU-baseline leg.  The product path never
imports it.  The oracle exports the pd_ ABI under the or_ prefix, so the
same peridem_amd.abi.Engine class drives it.
"""
import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_LIB = None


def build(force=False):
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB


def load():
    global _LIB
    if _LIB is None:
        build()
        from peridem_amd.abi import _bind
        lib = C.CDLL(LIB)
        lib.or_set_ranks.argtypes = [C.c_void_p, C.c_int]
        lib.or_set_ranks.restype = C.c_int
        lib.or_set_prune.argtypes = [C.c_void_p, C.c_int]
        lib.or_set_prune.restype = C.c_int
        lib.or_get_stretch.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_double), C.c_uint64,
                                       C.POINTER(C.c_uint64)]
        lib.or_get_stretch.restype = C.c_int
        lib.or_segment_mass3.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
        lib.or_segment_mass3.restype = None
        _LIB = (lib, _bind(lib, "or_"))
    return _LIB


def engine(system, ranks=1, prune=False):
    """prune: contact candidates from a cell grid (bitwise the same results,
    for large dense systems); off: the reference's all-pairs loops."""
    from peridem_amd.abi import Engine
    lib, fns = load()
    e = Engine(lib, fns, system, "or_")
    if ranks != 1:
        e._check(lib.or_set_ranks(e.ctx, ranks), "set_ranks")
    if prune:
        e._check(lib.or_set_prune(e.ctx, 1), "set_prune")
    return e


def stretch(e, p):
    import numpy as np
    lib, _ = load()
    m = C.c_uint64()
    e._check(lib.or_get_stretch(e.ctx, p, None, 0, C.byref(m)), "get_stretch")
    out = np.empty(m.value)
    e._check(lib.or_get_stretch(e.ctx, p, out.ctypes.data_as(C.POINTER(C.c_double)), m.value,
                                C.byref(m)), "get_stretch")
    return out


def segment_mass3(R, d):
    lib, _ = load()
    out = (C.c_double * 2)()
    lib.or_segment_mass3(R, d, out)
    return out[0], out[1]
