"""bench.py's roofline bookkeeping, on the CPU: the HBM traffic per launch comes
from a committed counter summary that is only used while it describes the
kernel sources and launch shape being timed."""
import csv
import os

import pytest

import bench

LAUNCH = {"waves_per_block": 1, "unroll": 16, "k_tile": 1, "xcd_remap": 1, "nontemporal": 1,
          "accumulate": 0, "rows_per_lane": 1, "blocks_per_cu": 0, "stage_lds": 7}


def _counter_dir(path, counter, kernel_values, copy_values):
    """A rocprofv3 --pmc output directory with one csv: the SpMV dispatches
    and the 1 GiB copy (plus a small upload copy, which must be ignored)."""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for v in kernel_values:
            w.writerow({"Kernel_Name": "void wk::spmv_winpk_kernel<PF64, false, 0, 2>(...)",
                        "Counter_Name": counter, "Counter_Value": v})
        for v in copy_values:
            w.writerow({"Kernel_Name": "__amd_rocclr_copyBuffer", "Counter_Name": counter,
                        "Counter_Value": v})
        w.writerow({"Kernel_Name": "__amd_rocclr_copyBuffer", "Counter_Name": counter,
                    "Counter_Value": 12.0})


def test_launch_key_names_every_kernel_field():
    key = bench.launch_key(LAUNCH)
    assert key.split() == [f"{k}={LAUNCH[k]}" for k in bench.KERNEL_FIELDS]


def test_write_then_read_traffic(tmp_path, monkeypatch):
    """Counts scale by what the same counter reads for the 1 GiB copy; the file
    is refused for other sources or another launch shape."""
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    monkeypatch.setattr(bench, "source_hash", lambda: "0123456789abcdef")
    os.makedirs(tmp_path / "profiles")
    fetch, write = str(tmp_path / "fetch"), str(tmp_path / "write")
    # reads count twice (FETCH_SIZE of a 1 GiB copy = 524288 KiB), writes once
    _counter_dir(fetch, "FETCH_SIZE", [1000.0, 1010.0, 990.0], [524288.0, 524288.0])
    _counter_dir(write, "WRITE_SIZE", [50.0, 50.0, 50.0], [1048576.0])
    bench.write_traffic("unit", fetch, write, LAUNCH)
    expected = (1000.0 * 2.0 + 50.0 * 1.0) * 1024
    got, src = bench.read_traffic("unit", LAUNCH)
    assert got == pytest.approx(expected, abs=1.0)
    assert src == os.path.join("profiles", "traffic_unit.txt")
    other = dict(LAUNCH, unroll=8)
    assert bench.read_traffic("unit", other) == (None, "traffic file measured with another launch shape")
    monkeypatch.setattr(bench, "source_hash", lambda: "fedcba9876543210")
    assert bench.read_traffic("unit", LAUNCH) == (None, "traffic file measured on other kernel sources")
    assert bench.read_traffic("none", LAUNCH) == (None, "no traffic file")


def test_committed_headline_traffic_matches_the_sources():
    """The committed headline traffic file was measured on the kernel sources in
    this tree with the headline's launch shape, so bench.py reports it."""
    got, src = bench.read_traffic("hex_elastic_3d_64", LAUNCH)
    assert got is not None, src
    assert 5e8 < got < 8e8  # bytes per launch: about 580 MB, below the algorithmic 790 MB
