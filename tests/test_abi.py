"""The C-ABI library builds, loads and exports every symbol its header declares."""
import os
import re
import subprocess
import ctypes as C

import pytest

from peridem_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "peridem_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pd_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("pd_create", "pd_add_particle", "pd_commit", "pd_step", "pd_get_state",
              "pd_get_connectivity", "pd_get_wall", "pd_destroy"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    assert os.path.exists(abi.GPU_LIB), "library not built"
    out = subprocess.run(["nm", "-D", "--defined-only", abi.GPU_LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_struct_layout_matches_header(tmp_path):
    """The ctypes structs mirror include/peridem_gpu.h: compare every field
    offset and the struct sizes with what the C compiler lays out."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "peridem_gpu.h"', "int main(){"]
    for cname, cls in (("pd_run_desc", abi.RunDesc), ("pd_particle_desc", abi.ParticleDesc),
                       ("pd_counters", abi.Counters), ("pd_pool_desc", abi.PoolDesc),
                       ("pd_profile", abi.Profile), ("pd_serve_event", abi.ServeEvent)):
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in filter(None, got):
        cname, f, v = line.split()
        cls = {"pd_run_desc": abi.RunDesc, "pd_particle_desc": abi.ParticleDesc,
               "pd_counters": abi.Counters, "pd_pool_desc": abi.PoolDesc,
               "pd_profile": abi.Profile, "pd_serve_event": abi.ServeEvent}[cname]
        want = C.sizeof(cls) if f == "size" else getattr(cls, f).offset
        assert int(v) == want, (cname, f, v, want)


def test_library_loads_without_gpu():
    lib, fns = abi.load_gpu_lib()
    for s in abi.SYMBOLS:
        assert s in fns


def test_bond_shape_table_ends_with_the_wide_shape():
    """The tuned shapes come first; the index one past them is the unstaged
    kernel with 64-bit record offsets; the index after that is no shape.
    Table queries touch no device."""
    shapes = abi.bond_shapes()
    k = abi.wide_bond_shape()
    assert k == len(shapes) and any(sh[2] for sh in shapes)
    assert abi.bond_shape_info(k) == (4, 2, False)
    with pytest.raises(ValueError):
        abi.bond_shape_info(k + 1)
    with pytest.raises(ValueError):
        abi.bond_shape_info(-1)


def test_connectivity_read_back_is_one_call_unless_the_bound_is_short():
    """The wrappers pass the particle's initial bond count as capacity and read
    the pairs in one call; a library that reports more pairs than that (the
    capacity check failing) is asked for the count first."""
    import ctypes as C
    import numpy as np
    from peridem_amd import abi
    pairs = np.array([[0, 1], [0, 2], [1, 3]], dtype=np.uint32)
    calls = []

    def fake(buf, cap, m):
        calls.append(cap)
        m._obj.value = len(pairs)
        if buf is None:
            return 0
        if cap < len(pairs):
            return -1
        C.memmove(buf, pairs.ctypes.data, pairs.nbytes)
        return 0

    def check(rc, what):
        assert rc == 0, what

    got = abi._read_connectivity(fake, 5, check)
    assert np.array_equal(got, pairs) and calls == [5]
    calls.clear()
    got = abi._read_connectivity(fake, 2, check)
    assert np.array_equal(got, pairs) and calls == [2, 0, 3]
