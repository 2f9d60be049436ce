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
                       ("pd_profile", abi.Profile)):
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
               "pd_profile": abi.Profile}[cname]
        want = C.sizeof(cls) if f == "size" else getattr(cls, f).offset
        assert int(v) == want, (cname, f, v, want)


def test_library_loads_without_gpu():
    lib, fns = abi.load_gpu_lib()
    for s in abi.SYMBOLS:
        assert s in fns
