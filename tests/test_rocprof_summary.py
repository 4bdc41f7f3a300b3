"""The rocprofv3 summariser on synthetic CSVs (no GPU): rows of cells, full-state
rows, calibrated bytes and the VALU figures are computed as documented."""
from __future__ import annotations

import csv

from ofp import rocprof_summary as rs

CELL = "void ofp::(anonymous namespace)::cell_kernel<14, unsigned int, 1, 2>(ofp::CellLaunch)"
CAPPED = "void ofp::(anonymous namespace)::cell_capped_kernel<14, unsigned int>(ofp::CellLaunch)"


def _trace(path, launches):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        t = 0
        for name, ms in launches:
            w.writerow({"Kernel_Name": name, "Start_Timestamp": t, "End_Timestamp": t + int(ms * 1e6)})
            t += int(ms * 1e6) + 1000


def test_trace_rows_and_full_state_mean(tmp_path, capsys):
    cells = 3
    # row 0 capped, rows 1 and 2 full-state
    launches = [(CAPPED, 1.0)] * cells + [(CELL, 10.0), (CELL, 11.0), (CELL, 12.0)] + [(CELL, 20.0)] * cells
    p = tmp_path / "t.csv"
    _trace(p, launches)
    rs.trace(str(p), cells)
    out = capsys.readouterr().out
    assert "KC=14: 3 rows of 3 cells" in out
    assert "3.0(3) 33.0(0) 60.0(0)" in out
    assert "median full-state row 60.0 ms over 2 full-state rows; per step (p: ms): 3:20.0 2:20.0 1:20.0" in out


PAIR = "void ofp::(anonymous namespace)::pair_kernel<14, unsigned int, {d}, 0, 2>(ofp::pair::PairLaunch)"


def test_trace_rows_with_pair_passes(tmp_path, capsys):
    """A pair pass (its launches d = -1, 0, 1, -2, 2; here one is skipped) counts as
    two cells; consecutive passes are told apart by the launch order."""
    cells = 5
    row = [(PAIR.format(d=d), ms) for d, ms in ((-1, 1.0), (0, 2.0), (1, 1.0), (-2, 0.5), (2, 0.5))]
    row += [(PAIR.format(d=d), ms) for d, ms in ((-1, 1.5), (0, 2.5), (1, 1.0), (2, 1.0))]
    row += [(CELL, 7.0)]
    p = tmp_path / "t.csv"
    _trace(p, [(CAPPED, 1.0)] * cells + row + row)
    rs.trace(str(p), cells)
    out = capsys.readouterr().out
    assert "KC=14: 3 rows of 5 cells" in out
    assert "5.0(5) 18.0(0) 18.0(0)" in out
    assert "per step (p: ms): 5,4:5.0 3,2:6.0 1:7.0" in out


def test_trace_pair_pass_with_overlapping_launches(tmp_path, capsys):
    """A pass whose readers run side by side (three streams) is timed by its wall
    span, not by the sum of its launches' durations."""
    ms = 1_000_000
    rows = []
    for base in (0, 100 * ms):  # two rows of one pass (2 cells) each
        rows += [(PAIR.format(d=-1), base + 0, base + 10 * ms), (PAIR.format(d=0), base + 0, base + 12 * ms),
                 (PAIR.format(d=1), base + 1 * ms, base + 9 * ms), (PAIR.format(d=-2), base + 12 * ms,
                 base + 15 * ms), (PAIR.format(d=2), base + 12 * ms, base + 14 * ms)]
    p = tmp_path / "t.csv"
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for name, t0, t1 in rows:
            w.writerow({"Kernel_Name": name, "Start_Timestamp": t0, "End_Timestamp": t1})
    rs.trace(str(p), 2)
    out = capsys.readouterr().out
    assert "KC=14: 2 rows of 2 cells" in out
    assert "15.0(0) 15.0(0)" in out
    assert "per step (p: ms): 2,1:15.0" in out
    rs.trace(str(p), 2, {1, 4})  # rows of other moduli counts only: none here
    assert "KC=14" not in capsys.readouterr().out


def test_counter_calibration(tmp_path, capsys):
    p = tmp_path / "c.csv"
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value", "Dispatch_Id"])
        w.writeheader()
        for d in range(4):
            w.writerow({"Kernel_Name": CELL, "Counter_Name": "FETCH_SIZE", "Counter_Value": 1000.0 * (d + 1),
                        "Dispatch_Id": d})
    rs.counter(str(p), 2.0, 2)
    out = capsys.readouterr().out
    # mean of 1000..4000 KiB = 2500 KiB, x1024 x2.0 bytes
    assert f"calibrated_GB={2500 * 1024 * 2.0 / 1e9:10.3f}" in out
    assert "KC=14: full-state rows 2" in out


def test_valu_busy_and_lanes(tmp_path, capsys):
    p = tmp_path / "v.csv"
    cells = 2
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value", "Dispatch_Id", "Start_Timestamp", "End_Timestamp"]
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for d in range(cells):
            vals = {"SQ_ACTIVE_INST_VALU": 256 * 1000.0, "SQ_THREAD_CYCLES_VALU": 256 * 1000.0 * 32,
                    "SQ_INSTS_VALU": 256 * 1000.0, "SQ_INSTS_SALU": 0.0, "SQ_INSTS_LDS": 0.0,
                    "SQ_INSTS_VMEM_RD": 0.0, "SQ_INSTS_VMEM_WR": 0.0, "SQ_WAVES": 64.0,
                    "GRBM_GUI_ACTIVE": 8 * 4000.0}
            for name, v in vals.items():
                w.writerow({"Kernel_Name": CELL, "Counter_Name": name, "Counter_Value": v, "Dispatch_Id": d,
                            "Start_Timestamp": 0, "End_Timestamp": 1000000})
    rs.valu(str(p), cells, 256, 8, None)
    out = capsys.readouterr().out
    # 1000 VALU issue cycles per CU out of 4000 per XCD: 25 %; half the lanes active: 50 %
    assert "  25.0   50.0" in out
    assert "row: VALU busy 25.0%" in out


def test_valu_sums_a_pair_pass(tmp_path, capsys):
    """The launches of one pair pass (d = -1, 0, 1, -2, 2) are one step of two
    cells, their counters summed."""
    p = tmp_path / "v.csv"
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value", "Dispatch_Id", "Start_Timestamp", "End_Timestamp"]
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        did = 0
        for row in range(2):
            for d, busy in ((-1, 1000.0), (0, 3000.0), (1, 1000.0), (-2, 1000.0), (2, 2000.0)):
                vals = {"SQ_ACTIVE_INST_VALU": 256 * busy, "SQ_THREAD_CYCLES_VALU": 256 * busy * 64,
                        "SQ_INSTS_VALU": 0.0, "SQ_INSTS_SALU": 0.0, "SQ_INSTS_LDS": 0.0,
                        "SQ_INSTS_VMEM_RD": 0.0, "SQ_INSTS_VMEM_WR": 0.0, "SQ_WAVES": 64.0,
                        "GRBM_GUI_ACTIVE": 8 * 4000.0}
                for name, v in vals.items():
                    w.writerow({"Kernel_Name": PAIR.format(d=d), "Counter_Name": name, "Counter_Value": v,
                                "Dispatch_Id": did, "Start_Timestamp": 0, "End_Timestamp": 1000000})
                did += 1
    rs.valu(str(p), 2, 256, 8, None)
    out = capsys.readouterr().out
    assert "(2 full-state rows)" in out
    # 8000 VALU issue cycles per CU over 5 x 4000 cycles: 40 %, every lane active
    assert "  2,1    5.00   40.0  100.0" in out


SORTED = "void ofp::(anonymous namespace)::cell_sorted_kernel<14, unsigned int>(ofp::CellLaunch)"


def test_dram_bytes_per_family(tmp_path, capsys):
    """DRAM request counters are in 32-byte units: bytes = 32 x count, averaged per
    dispatch of each cell-kernel family; other kernels are left out."""
    p = tmp_path / "d.csv"
    rows = [(1, SORTED, 1000, 500), (2, SORTED, 3000, 700), (3, CELL, 10, 20),
            (4, "void ofp::set_one_values_kernel(unsigned int*)", 1, 1)]
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for d, name, rd, wr in rows:
            w.writerow({"Dispatch_Id": d, "Kernel_Name": name, "Counter_Name": rs.DRAM_COUNTERS[0],
                        "Counter_Value": rd})
            w.writerow({"Dispatch_Id": d, "Kernel_Name": name, "Counter_Name": rs.DRAM_COUNTERS[1],
                        "Counter_Value": wr})
    got = rs.dram_bytes(str(p))
    assert set(got) == {"cell_sorted_kernel<14>", "cell_kernel<14>"}
    assert got["cell_sorted_kernel<14>"] == {"read": 32.0 * 2000, "write": 32.0 * 600, "dispatches": 2,
                                             "read_total": 32.0 * 4000, "write_total": 32.0 * 1200}
    assert got["cell_kernel<14>"] == {"read": 320.0, "write": 640.0, "dispatches": 1,
                                      "read_total": 320.0, "write_total": 640.0}
    rs.dram(str(p))
    assert "cell_sorted_kernel<14>" in capsys.readouterr().out


def test_kernel_family_names():
    assert rs.kernel_family(SORTED) == "cell_sorted_kernel<14>"
    assert rs.kernel_family(CELL) == "cell_kernel<14>"
    assert rs.kernel_family(CAPPED) == "cell_capped_kernel<14>"
    assert rs.kernel_family("void ofp::fill_remote_kernel()") is None
    assert rs.kernel_family("void ofp::(anonymous namespace)::pair_kernel<14, unsigned int, -1>"
                            "(ofp::pair::PairLaunch)") == "pair_kernel<14>"


def test_launches_groups_pair_dispatches_into_passes(tmp_path, capsys):
    """Per-dispatch listing: DRAM bytes = counter x 32, pair launches grouped into
    passes where the d order restarts, other kernels listed alone."""
    p = tmp_path / "l.csv"
    M = 10**7
    seq = [(PAIR.format(d=-1), 2.0, 100 * M, 50 * M), (PAIR.format(d=0), 3.0, 200 * M, 100 * M),
           (PAIR.format(d=2), 1.0, 10 * M, 10 * M), (PAIR.format(d=-1), 2.0, 300 * M, 30 * M),
           (PAIR.format(d=-2), 1.0, 40 * M, 20 * M), (CELL, 4.0, 1000 * M, 500 * M)]
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value",
                                          "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        t = 0
        for i, (name, ms, rd, wr) in enumerate(seq):
            for cname, v in (("TCC_EA0_RDREQ_DRAM_32B_sum", rd), ("TCC_EA0_WRREQ_WRITE_DRAM_32B_sum", wr)):
                w.writerow({"Dispatch_Id": i + 1, "Kernel_Name": name, "Counter_Name": cname, "Counter_Value": v,
                            "Start_Timestamp": t, "End_Timestamp": t + int(ms * 1e6)})
            t += int(ms * 1e6) + 1000
    rs.launches(str(p))
    lines = capsys.readouterr().out.splitlines()
    totals = [ln for ln in lines if "pass total" in ln]
    assert len(totals) == 2
    # first pass: 6 ms, (100 + 200 + 10) M x 32 B = 99.2 GB read, 160 M x 32 B = 51.2 GB written
    assert totals[0].split()[-4:-1] == ["6.00", "99.2", "51.2"]
    assert totals[1].split()[-4] == "3.00"
    assert any(ln.split()[1] == "cell_kernel<14>" and ln.split()[2] == "4.00" for ln in lines if ln.strip())
