"""The rocprofv3 summary helper (mpm_amd/profiling.py) on synthetic counter
files: per-kernel means, the kernel filter and the SQ ratios."""
import csv

from mpm_amd.profiling import summarize_csv

COLS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=COLS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(COLS, r)))


def test_means_ratios_and_filter(tmp_path):
    a, b = tmp_path / "a.csv", tmp_path / "b.csv"
    _write(a, [(1, "void k_fused<1>(mpmk::ItemDesc)", "SQ_LDS_BANK_CONFLICT", 30.0),
               (2, "void k_fused<1>(mpmk::ItemDesc)", "SQ_LDS_BANK_CONFLICT", 50.0),
               (3, "k_other(int)", "SQ_WAVES", 7.0)])
    _write(b, [(1, "void k_fused<1>(mpmk::ItemDesc)", "SQ_ACTIVE_INST_LDS", 80.0),
               (1, "void k_fused<1>(mpmk::ItemDesc)", "SQ_WAVES", 4.0),
               (1, "void k_fused<1>(mpmk::ItemDesc)", "SQ_INSTS_VALU", 10.0)])
    out = summarize_csv([a, b])
    lines = out.splitlines()
    assert "void k_fused<1>" in lines and "k_other" in lines
    conflict = next(l for l in lines if "SQ_LDS_BANK_CONFLICT" in l)
    assert "40.0" in conflict and "mean of 2 dispatches" in conflict
    ratio = next(l for l in lines if "lds_bank_conflict_per_issue" in l)
    assert float(ratio.split()[2]) == 0.5
    valu = next(l for l in lines if "valu_insts_per_wave" in l)
    assert float(valu.split()[2]) == 2.5
    # a ratio needs both counters: none for the kernel that has only SQ_WAVES
    assert out.count("= ") == 2
    only = summarize_csv([a, b], kernels="fused")
    assert "k_other" not in only and "k_fused" in only


def test_busy_line_counts_gaps_not_overlaps(tmp_path):
    import sqlite3

    from mpm_amd.profiling import busy_line, summarize_db
    # two streams: [0, 10) and [5, 20) overlap, a gap to [25, 30), then [30, 40)
    rows = [("void k_a<1>(int)", 0, 10_000), ("k_b(int)", 5_000, 20_000),
            ("void k_a<1>(int)", 25_000, 30_000), ("k_b(int)", 30_000, 40_000)]
    line = busy_line(rows)
    assert "4 launches over 40.0 us" in line and "running 35.0 us" in line and "none 5.0 us (12.50%)" in line
    db = tmp_path / "run_results.db"
    con = sqlite3.connect(str(db))
    con.execute("create table kernels (name text, start integer, end integer)")
    con.executemany("insert into kernels values (?, ?, ?)", rows + [("k_after(int)", 100_000, 110_000)])
    con.commit()
    con.close()
    # from the 2nd last k_a on (at 0 us), the launches starting before the last k_a ends (30 us)
    out = summarize_db(db, tail_of="void k_a<1>:2", busy=True)
    assert "busy: 3 launches over 30.0 us" in out and "none 5.0 us (16.67%)" in out
    assert "k_after" in out  # the table keeps every launch from the start of the tail
    out = summarize_db(db, tail_of="void k_a<1>:1", busy=True)
    assert "busy: 1 launches over 5.0 us" in out and "none 0.0 us" in out
    out = summarize_db(db, busy=True)
    assert "5 launches over 110.0 us" in out
