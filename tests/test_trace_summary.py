"""The profile summariser (trace_summary.py) on small synthetic rocprofv3 CSVs:
per-kernel statistics, the last-N selection of a kernel trace, and the
counter report's copy calibration, idle filter, grid filter and per-64-point
scaling."""
import csv

import pytest

from multigpu_advectiondiffusion_amd import trace_summary as T


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_kernel_trace_last_n(tmp_path, capsys):
    p = tmp_path / "k_kernel_trace.csv"
    rows = []
    t = 0
    for i in range(6):  # stage kernel 100, 200, ..., 600 ns; one copy in between
        rows.append([i, "stage", t, t + 100 * (i + 1)])
        t += 1000
        if i == 2:
            rows.append([100 + i, "copy", t, t + 50])
            t += 1000
    _write(p, ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"], rows)
    disp = T.read_dispatches(str(p))
    assert [n for n, _, _ in disp].count("stage") == 6
    assert T.main([str(p), "--kernel", "stage", "--last", "2"]) == 0
    out = capsys.readouterr().out
    # the last two stage dispatches last 500 and 600 ns
    assert "selected: 2 dispatches, average 550 ns" in out
    assert "| `stage` | 6 | 2100 | 350 |" in out


def _pmc_rows():
    hdr = ["Dispatch_Id", "Grid_Size", "Kernel_Name", "Workgroup_Size", "VGPR_Count", "Accum_VGPR_Count",
           "LDS_Block_Size", "Scratch_Size", "Counter_Name", "Counter_Value"]
    rows = []
    did = 0
    # three stage dispatches of 1024 threads, one idle (after t_end), one of another size
    for grid, valu, fetch in ((1024, 6400.0, 1000.0), (1024, 6600.0, 1100.0), (1024, 10.0, 1.0), (512, 999.0, 5.0)):
        did += 1
        rows.append([did, grid, "stage", 256, 64, 0, 0, 0, "SQ_INSTS_VALU", valu])
        rows.append([did, grid, "stage", 256, 64, 0, 0, 0, "FETCH_SIZE", fetch])
    for _ in range(3):  # calibration copies of 64 KiB reading half of it on the counter
        did += 1
        rows.append([did, 4096, "__amd_rocclr_copyBuffer", 256, 8, 0, 0, 0, "FETCH_SIZE", 32.0])
    return hdr, rows


def test_pmc_report_scaling_and_filters(tmp_path):
    p = tmp_path / "c_counter_collection.csv"
    _write(p, *_pmc_rows())
    rep = T.pmc_report([str(p)], "stage", points=4096, calib_kib=64.0, calib_n=3, skip_idle=True, grid=1024)
    assert "FETCH_SIZE: last 3 copies of 64 KiB -> factor 2.00000" in rep
    # idle dispatch (10 < 5 % of 6600) dropped: mean of 6400 and 6600, per 64 points of 4096
    assert "SQ_INSTS_VALU: mean 6500 per dispatch" in rep
    assert "= 101.6 per 64 points" in rep
    # fetch mean 1050 KiB x factor 2 x 1024 B
    assert "-> 2.1504e+06 bytes" in rep
    assert "grid 1024 threads" in rep
    rep_all = T.pmc_report([str(p)], "stage")
    assert "n 4)" in rep_all  # no filters: all four stage dispatches


def test_pmc_cli(tmp_path, capsys):
    p = tmp_path / "c_counter_collection.csv"
    _write(p, *_pmc_rows())
    assert T.main(["--pmc", str(p), "--kernel", "stage", "--grid", "512", "--points", "2048"]) == 0
    out = capsys.readouterr().out
    assert "SQ_INSTS_VALU: mean 999 per dispatch" in out
    with pytest.raises(SystemExit):
        T.main([str(p), str(p)])  # a kernel trace summary reads one file
