"""The rocprofv3 summariser: per-launch means leave out the autotuner's
dispatches, and with several passes each kernel's figure comes from the pass
in which it ran most (its step launches, not the tuner's back-to-back ones)."""
import csv

import pytest

from peridem_amd import profsum

COPY = "__amd_rocclr_copyBuffer"
K3 = "void pd::k_bond_lds<3, 2>(pd::Dev)"
K4 = "void pd::k_bond_lds<4, 2>(pd::Dev)"


def write_pass(path, counter, launches):
    """launches: [(kernel, value KiB)] in dispatch order."""
    path.mkdir()
    with open(path / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, ["Kernel_Name", "Counter_Name", "Dispatch_Id", "Counter_Value"])
        w.writeheader()
        for i, (k, v) in enumerate(launches):
            w.writerow({"Kernel_Name": k, "Counter_Name": counter, "Dispatch_Id": i + 1, "Counter_Value": v})
    return str(path)


def pinned(tmp_path, name, counter, steady, tuned, copy_kib):
    """A pass pinned to `steady` (10 dispatches of 300 KiB after 4 of 100),
    with 4 tuner launches of `tuned` (100 KiB) and the calibrating copy."""
    launches = [(steady, 100.0)] * 4 + [(tuned, 100.0)] * 4 + [(steady, 300.0)] * 10 + [(COPY, copy_kib)]
    return write_pass(tmp_path / name, counter, launches)


def test_traffic_takes_each_kernel_from_the_pass_where_it_ran_most(tmp_path):
    mib = 1
    # FETCH_SIZE reads half the copied bytes (scale 2), WRITE_SIZE all of them
    f3 = pinned(tmp_path, "f3", "FETCH_SIZE", K3, K4, 512.0)
    f4 = pinned(tmp_path, "f4", "FETCH_SIZE", K4, K3, 512.0)
    w3 = pinned(tmp_path, "w3", "WRITE_SIZE", K3, K4, 1024.0)
    w4 = pinned(tmp_path, "w4", "WRITE_SIZE", K4, K3, 1024.0)
    t = profsum.traffic(f"{f3},{f4}", f"{w3},{w4}", mib, 10, 20)
    assert t["fetch_scale"] == pytest.approx(2.0) and t["write_scale"] == pytest.approx(1.0)
    # steady launches only: 300 KiB fetched (x2) + 300 KiB written
    for k in ("pd::k_bond_lds<3, 2>", "pd::k_bond_lds<4, 2>"):
        assert t["per_kernel_bytes"][k] == pytest.approx((2 * 300 + 300) * 1024)
    assert t["nodes"] == 10 and t["movable_bonds"] == 20


def test_single_pass_keeps_the_tuner_launches_of_shapes_it_only_timed(tmp_path):
    f3 = pinned(tmp_path, "f3", "FETCH_SIZE", K3, K4, 512.0)
    w3 = pinned(tmp_path, "w3", "WRITE_SIZE", K3, K4, 1024.0)
    t = profsum.traffic(f3, w3, 1, 10, 20)
    assert t["per_kernel_bytes"]["pd::k_bond_lds<4, 2>"] == pytest.approx((2 * 100 + 100) * 1024)
    assert t["bond_kernel"] == "pd::k_bond_lds<3, 2>"


def test_fp64_merges_passes(tmp_path):
    a = tmp_path / "a"
    a.mkdir()
    rows = []
    for i in range(8):
        for c, v in zip(profsum.FP64_COUNTERS, (1.0, 2.0, 3.0, 4.0)):
            rows.append({"Kernel_Name": K3, "Counter_Name": c, "Dispatch_Id": i + 1, "Counter_Value": v})
    with open(a / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, ["Kernel_Name", "Counter_Name", "Dispatch_Id", "Counter_Value"])
        w.writeheader()
        w.writerows(rows)
    b = tmp_path / "b"
    b.mkdir()
    with open(b / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, ["Kernel_Name", "Counter_Name", "Dispatch_Id", "Counter_Value"])
        w.writeheader()
        for c, v in zip(profsum.FP64_COUNTERS, (9.0, 9.0, 9.0, 9.0)):
            w.writerow({"Kernel_Name": K4, "Counter_Name": c, "Dispatch_Id": 1, "Counter_Value": v})
    t = profsum.fp64(f"{a},{b}")["per_kernel_fp64_wave_instr"]
    assert t["pd::k_bond_lds<3, 2>"] == pytest.approx(10.0)
    assert t["pd::k_bond_lds<4, 2>"] == pytest.approx(36.0)
