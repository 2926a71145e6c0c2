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
