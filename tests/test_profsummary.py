"""The rocprofv3 CSV summary behind the committed roofline files: averages,
the per-shape grid filter and the copy-kernel calibration of the HBM
counters, on small synthetic CSVs."""

import csv
import json

import pytest

from latticeboltzmann_amd import profsummary as ps

K = "void lbm::k_sweep<double, 1, false, 5, true, false>(double const*)"
OTHER = "void lbm::k_relayout<double>(double const*)"


def _write(path, header, rows):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(header)
        w.writerows(rows)
    return str(path)


def _trace(tmp_path):
    # two launch shapes of one kernel (grids 1024 and 2048 work-items) and another kernel
    hdr = ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X", "Grid_Size_Y",
           "Grid_Size_Z"]
    rows = [[K, 0, 1000, 1024, 1, 1], [K, 2000, 5000, 1024, 1, 1], [K, 6000, 16000, 2048, 1, 1],
            [OTHER, 20000, 20500, 4096, 1, 1]]
    return _write(tmp_path / "kt_kernel_trace.csv", hdr, rows)


def _counters(tmp_path, name, values):
    hdr = ["Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]
    return _write(tmp_path / name, hdr, values)


def test_trace_average_by_grid(tmp_path):
    tr = _trace(tmp_path)
    assert ps.trace_average_us(tr, "k_sweep")[1:] == (3, 14000 / 3 / 1e3)
    assert ps.trace_average_us(tr, "k_sweep", 1024)[1:] == (2, 2.0)
    assert ps.trace_average_us(tr, "k_sweep", 2048)[1:] == (1, 10.0)
    with pytest.raises(KeyError):
        ps.trace_average_us(tr, "k_sweep", 4096)


def test_stats_csv_has_no_grid(tmp_path):
    st = _write(tmp_path / "kt_kernel_stats.csv",
                ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs"],
                [[K, 3, 14000, 14000 / 3, 96.6, 1000, 10000]])
    assert ps.trace_average_us(st, "k_sweep")[1] == 3
    with pytest.raises(ValueError):
        ps.trace_average_us(st, "k_sweep", 1024)


def test_summary_scales_hbm_counters_by_the_copy_calibration(tmp_path):
    tr = _trace(tmp_path)
    pmc = _counters(tmp_path, "p.csv", [
        [1024, K, "FETCH_SIZE", 100.0], [1024, K, "FETCH_SIZE", 300.0], [2048, K, "FETCH_SIZE", 9e9],
        [1024, K, "WRITE_SIZE", 50.0], [2048, K, "WRITE_SIZE", 9e9],
        [1024, K, "SQ_INSTS_VALU", 7.0]])
    # the copy kernel moves 2 KiB each way; its FETCH_SIZE reads half of that, WRITE_SIZE all of it
    cf = _counters(tmp_path, "cf.csv", [[4096, OTHER, "FETCH_SIZE", 1.0]])
    cw = _counters(tmp_path, "cw.csv", [[4096, OTHER, "WRITE_SIZE", 2.0]])
    out = ps.summarize("k_sweep", tr, [pmc], cf, cw, 2048, grid=1024)
    assert out["grid_work_items"] == 1024 and out["kernel_trace_calls"] == 2
    assert out["counters_per_dispatch"]["FETCH_SIZE"] == 200.0
    assert out["valu_wave_insts_per_launch"] == 7.0
    assert out["fetch_scale"] == 2.0 and out["write_scale"] == 1.0
    assert out["hbm_read_bytes_per_launch"] == 200 * 1024 * 2
    assert out["hbm_write_bytes_per_launch"] == 50 * 1024
    assert out["hbm_bytes_per_launch"] == 200 * 1024 * 2 + 50 * 1024


def test_cli_merges_entries(tmp_path):
    tr = _trace(tmp_path)
    pmc = _counters(tmp_path, "p.csv", [[1024, K, "SQ_WAVES", 16.0], [2048, K, "SQ_WAVES", 32.0]])
    out = str(tmp_path / "roofline.json")
    ps.main(["--kernel", "k_sweep", "--key", "a", "--trace", tr, "--pmc", pmc, "--grid", "1024",
             "--out", out])
    ps.main(["--kernel", "k_sweep", "--key", "b", "--trace", tr, "--pmc", pmc, "--grid", "2048",
             "--out", out, "--merge"])
    doc = json.load(open(out))["per_variant"]
    assert doc["a"]["counters_per_dispatch"]["SQ_WAVES"] == 16.0
    assert doc["b"]["counters_per_dispatch"]["SQ_WAVES"] == 32.0
    assert "hbm_bytes_per_launch" not in doc["a"]  # no calibration given
