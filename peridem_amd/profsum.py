"""Summaries of rocprofv3 CSV output for the profiles/ folder.

  python -m peridem_amd.profsum trace <dir>            kernel-trace --stats table
  python -m peridem_amd.profsum steps <dir>            the same from the per-dispatch
        trace, each kernel's first SKIP dispatches left out when it has more
  python -m peridem_amd.profsum counters <dir> [kernel] per-dispatch PMC averages
  python -m peridem_amd.profsum traffic <fetch_dir> <write_dir> <copy_mib> <nodes> <bonds>
        HBM bytes per launch of every kernel, FETCH_SIZE/WRITE_SIZE scaled by
        what the counters read for the plain device copy of copy_mib MiB
  python -m peridem_amd.profsum fp64 <dir> [out.json]
        FP64 wave-instructions per launch of every kernel (SQ_INSTS_VALU_ADD/
        MUL/FMA/TRANS_F64), for the FP64 issue utilisation

Per-launch means (steps, counters, traffic, fp64) leave out each kernel's
first SKIP = 4 dispatches when it has more: the first step of a pool times
every bond-kernel shape with one warm-up and three timed launches on scratch
copies, without the contact terms and velocity half that the step's own
launches add, so the selected shape's step launches are the ones after those.
"""
import collections
import csv
import glob
import json
import sys


def _rows(d, name):
    path = glob.glob(f"{d}/**/*{name}", recursive=True)
    if not path:
        raise SystemExit(f"no {name} under {d}")
    with open(path[0]) as f:
        return list(csv.DictReader(f))


def _short(k):
    k = k.split("(")[0]
    return k[5:] if k.startswith("void ") else k


def trace(d):
    rows = _rows(d, "kernel_stats.csv")
    print("| kernel | calls | avg us | min us | max us | % of GPU time |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        print(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
              f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")


SKIP = 4


def _steps_only(vals):
    """Values in dispatch order, less the autotuner's launches."""
    vals = [v for _, v in sorted(vals)]
    return vals[SKIP:] if len(vals) > SKIP else vals


def _agg(d):
    raw = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in _rows(d, "counter_collection.csv"):
        raw[_short(r["Kernel_Name"])][r["Counter_Name"]].append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for k, cs in raw.items():
        for c, v in cs.items():
            out[k][c] = _steps_only(v)
    return out


def steps(d):
    per = collections.defaultdict(list)
    for r in _rows(d, "kernel_trace.csv"):
        per[_short(r["Kernel_Name"])].append(
            (int(r["Dispatch_Id"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3))
    rows = sorted(((k, _steps_only(v)) for k, v in per.items()), key=lambda kv: -sum(kv[1]))
    total = sum(sum(v) for _, v in rows)
    print("| kernel | launches | avg us | min us | max us | % of GPU time |")
    print("|---|---|---|---|---|---|")
    for k, v in rows:
        print(f"| `{k}` | {len(v)} | {sum(v) / len(v):.1f} | {min(v):.1f} | {max(v):.1f} | "
              f"{100 * sum(v) / total:.1f} |")


def counters(d, kernel=None):
    a = _agg(d)
    for k, cs in a.items():
        if kernel and kernel not in k:
            continue
        print(f"`{k}` ({len(next(iter(cs.values())))} dispatches, mean per dispatch)")
        for c, v in sorted(cs.items()):
            print(f"  {c:32s} {sum(v) / len(v):18.1f}")


def _steady(dirs, counter):
    """Per kernel, the dispatches of the pass (of several, comma-separated
    directories) in which it ran most: a pass that pins the bond-kernel shape
    gives that shape's step launches, not the autotuner's back-to-back ones,
    which find their inputs still cached."""
    out = {}
    for d in dirs.split(","):
        for k, cs in _agg(d).items():
            if counter in cs and len(cs[counter]) > len(out.get(k, {}).get(counter, [])):
                out[k] = cs
    return out


def traffic(fd, wd, copy_mib, nodes, bonds):
    f, w = _steady(fd, "FETCH_SIZE"), _steady(wd, "WRITE_SIZE")
    copy_bytes = copy_mib * 1024 * 1024
    # the plain copy: the largest copyBuffer dispatch
    cf = max(f["__amd_rocclr_copyBuffer"]["FETCH_SIZE"])
    cw = max(w["__amd_rocclr_copyBuffer"]["WRITE_SIZE"])
    fs, ws = copy_bytes / (cf * 1024), copy_bytes / (cw * 1024)
    print(f"copy probe {copy_mib} MiB: FETCH_SIZE {cf:.0f} KiB -> scale {fs:.4f}; "
          f"WRITE_SIZE {cw:.0f} KiB -> scale {ws:.4f}")
    print("| kernel | fetched MB | written MB | HBM MB per launch |")
    print("|---|---|---|---|")
    res = {}
    for k in f:
        if not k.startswith("pd::") or k not in w:
            continue
        fb = sum(f[k]["FETCH_SIZE"]) / len(f[k]["FETCH_SIZE"]) * 1024 * fs
        wb = sum(w[k]["WRITE_SIZE"]) / len(w[k]["WRITE_SIZE"]) * 1024 * ws
        res[k] = fb + wb
        print(f"| `{k}` | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {(fb + wb) / 1e6:.1f} |")
    # the bond kernel: the shape with the most dispatches (the autotuner
    # times every shape a few times before the run settles on one)
    bond = [k for k in f if k.startswith("pd::k_bond")]
    bond = max(bond, key=lambda k: len(f[k]["FETCH_SIZE"])) if bond else None
    return {"nodes": nodes, "movable_bonds": bonds, "fetch_scale": fs, "write_scale": ws,
            "bond_kernel": bond, "hbm_bytes_per_launch": res.get(bond),
            "per_kernel_bytes": res}


FP64_COUNTERS = ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64",
                 "SQ_INSTS_VALU_TRANS_F64")


def fp64(d):
    a = _steady(d, FP64_COUNTERS[0])
    print("| kernel | ADD | MUL | FMA | TRANS | FP64 wave-instructions per launch | all VALU |")
    print("|---|---|---|---|---|---|---|")
    res = {}
    for k, cs in a.items():
        if not k.startswith("pd::") or FP64_COUNTERS[0] not in cs:
            continue
        m = {c: sum(cs[c]) / len(cs[c]) for c in FP64_COUNTERS}
        tot = sum(m.values())
        valu = sum(cs["SQ_INSTS_VALU"]) / len(cs["SQ_INSTS_VALU"]) if "SQ_INSTS_VALU" in cs else float("nan")
        res[k] = tot
        print(f"| `{k}` | {m[FP64_COUNTERS[0]]:.0f} | {m[FP64_COUNTERS[1]]:.0f} | {m[FP64_COUNTERS[2]]:.0f} | "
              f"{m[FP64_COUNTERS[3]]:.0f} | {tot:.0f} | {valu:.0f} |")
    return {"per_kernel_fp64_wave_instr": res}


if __name__ == "__main__":
    cmd = sys.argv[1]
    if cmd == "trace":
        trace(sys.argv[2])
    elif cmd == "steps":
        steps(sys.argv[2])
    elif cmd == "counters":
        counters(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else None)
    elif cmd == "traffic":
        t = traffic(sys.argv[2], sys.argv[3], int(sys.argv[4]), int(sys.argv[5]), int(sys.argv[6]))
        if len(sys.argv) > 7:
            with open(sys.argv[7], "w") as fh:
                json.dump(t, fh, indent=1)
    elif cmd == "fp64":
        t = fp64(sys.argv[2])
        if len(sys.argv) > 3:
            with open(sys.argv[3], "w") as fh:
                json.dump(t, fh, indent=1)
