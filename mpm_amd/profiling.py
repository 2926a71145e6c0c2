"""Profiling helpers for rocprofv3 runs of the engine.

Subcommands (run as ``python -m mpm_amd.profiling <cmd>``):

  run        advance one of the benchmark scenes for --steps substeps after
             --warmup substeps (the program profiled under rocprofv3)
  mesh       build the surface mesh of a benchmark scene --reps times after
             --warmup substeps (mpm_mesh_build at --side voxels, --radius)
  copy       a plain device-to-device copy of --mib MiB, --reps times (the
             calibration kernel for the FETCH_SIZE / WRITE_SIZE counters)
  summarize  per-kernel table of a rocprofv3 output: a rocpd database
             (kernel trace: calls, mean/min/max duration) or counter CSV files
             (mean value per dispatch of every counter, and the ratios of
             the SQ counters: LDS bank conflicts, instructions per wavefront);
             with --markers also the roctx ranges of a --marker-trace run;
             with --busy the time the device ran no kernel between the first
             and the last launch kept (the gaps between launches)

The first two need a GPU; `summarize` runs anywhere.
"""
from __future__ import annotations

import argparse
import csv
import re
import sqlite3
import sys
from collections import defaultdict
from pathlib import Path


def _clean(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*", "", name).strip()


def cmd_run(a) -> int:
    import torch
    from . import Engine, scenes
    sc = getattr(scenes, a.workload)()
    st = sc.state()
    eng = Engine(sc.params)
    eng.set_materials(sc.materials)
    for kv in a.option or []:
        k, v = kv.split("=")
        eng.set_option(k, int(v))
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    eng.advance(a.warmup)
    torch.cuda.synchronize()
    eng.advance(a.steps)
    torch.cuda.synchronize()
    eng.close()
    print(f"ran {a.workload}: {sc.n} particles, N={sc.params.N}, {a.warmup}+{a.steps} substeps",
          flush=True)
    return 0


def cmd_mesh(a) -> int:
    import time
    import torch
    from . import Engine, scenes
    sc = getattr(scenes, a.workload)()
    st = sc.state()
    eng = Engine(sc.params)
    eng.set_materials(sc.materials)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    eng.advance(a.warmup)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        t0 = time.perf_counter()
        m = eng.mesh(a.side, a.radius)
        t1 = time.perf_counter()
        print(f"mesh {a.workload} side={a.side} radius={a.radius}: {len(m['triangles'])} triangles, "
              f"{1e3 * (t1 - t0):.1f} ms (build + copy to host)", flush=True)
    eng.close()
    return 0


def cmd_copy(a) -> int:
    import torch
    n = a.mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB x {a.reps}", flush=True)
    return 0


def busy_line(rows) -> str:
    """Span from the first start to the last end of the launches (name,
    start, end in ns), the time some kernel ran (the union of their intervals:
    launches on several streams overlap) and the rest, when none ran."""
    iv = sorted((s, e) for _, s, e in rows)
    if not iv:
        return "busy: no launches"
    busy, cur_s, cur_e = 0, iv[0][0], iv[0][1]
    for s, e in iv[1:]:
        if s > cur_e:
            busy += cur_e - cur_s
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    busy += cur_e - cur_s
    span = max(e for _, e in iv) - iv[0][0]
    idle = span - busy
    return (f"busy: {len(iv)} launches over {span / 1e3:.1f} us, some kernel running {busy / 1e3:.1f} us, "
            f"none {idle / 1e3:.1f} us ({100.0 * idle / span if span else 0.0:.2f}%)")


def summarize_db(path: Path, tail_of: str = "", busy: bool = False) -> str:
    """Per-kernel durations. tail_of = "KERNEL:N" keeps only the launches from
    the N-th last launch of KERNEL on (the timed region of a run whose warm-up
    launches other configurations, e.g. autotuning trials); with it, busy
    keeps the launches up to the last launch of KERNEL for the busy line."""
    con = sqlite3.connect(str(path))
    rows = con.execute("select name, start, end from kernels").fetchall()
    head = ""
    if tail_of:
        kname, n = tail_of.rsplit(":", 1)
        starts = sorted(s for nm, s, _ in rows if _clean(nm) == kname)
        if len(starts) < int(n):
            raise SystemExit(f"{kname}: {len(starts)} launches, fewer than {n}")
        t0 = starts[-int(n)]
        rows = [r for r in rows if r[1] >= t0]
        head = f"(launches from the {n}-th last {kname} on)\n"
        if busy:
            t1 = max(e for nm, _, e in rows if _clean(nm) == kname)
            head += busy_line([r for r in rows if r[1] < t1]) + "\n"
    elif busy:
        head = busy_line(rows) + "\n"
    per = defaultdict(list)
    for n, s, e in rows:
        per[_clean(n)].append((e - s) / 1000.0)
    tot = sum(sum(v) for v in per.values()) or 1.0
    out = [f"{'kernel':44s} {'calls':>6s} {'total_us':>11s} {'mean_us':>9s} {'min_us':>9s} "
           f"{'max_us':>9s} {'pct':>6s}"]
    for n, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"{n[:44]:44s} {len(v):6d} {sum(v):11.1f} {sum(v) / len(v):9.2f} {min(v):9.2f} "
                   f"{max(v):9.2f} {100 * sum(v) / tot:6.2f}")
    return head + "\n".join(out)


def summarize_markers(path: Path) -> str:
    """roctx ranges of a --marker-trace run (engine option "markers"): count,
    total and mean host-side duration per range name."""
    import json
    con = sqlite3.connect(str(path))
    per = defaultdict(list)
    for ext, start, end in con.execute("select extdata, start, end from regions"):
        try:
            msg = json.loads(ext).get("message") if ext else None
        except ValueError:
            msg = None
        if msg:
            per[msg].append((end - start) / 1e3)
    out = [f"{'range':28s} {'calls':>6s} {'total_us':>11s} {'mean_us':>9s}   (host-side enqueue time)"]
    for n, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"{n[:28]:28s} {len(v):6d} {sum(v):11.1f} {sum(v) / len(v):9.2f}")
    return "\n".join(out)


# ratios printed under a kernel's counters when both counters were collected:
# (name, numerator, denominator, meaning)
_RATIOS = (
    ("lds_bank_conflict_per_issue", "SQ_LDS_BANK_CONFLICT", "SQ_ACTIVE_INST_LDS",
     "bank-conflict stall cycles per cycle spent issuing LDS instructions"),
    ("lds_addr_conflict_per_issue", "SQ_LDS_ADDR_CONFLICT", "SQ_ACTIVE_INST_LDS",
     "address-conflict stall cycles per LDS issue cycle"),
    ("valu_insts_per_wave", "SQ_INSTS_VALU", "SQ_WAVES", "vector ALU instructions per wavefront"),
    ("lds_insts_per_wave", "SQ_INSTS_LDS", "SQ_WAVES", "LDS instructions per wavefront"),
    ("salu_insts_per_wave", "SQ_INSTS_SALU", "SQ_WAVES", "scalar ALU instructions per wavefront"),
    ("lds_wait_frac", "SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES", "wave-cycles waiting on LDS / wave-cycles"),
    ("any_wait_frac", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES", "wave-cycles waiting on anything / wave-cycles"),
)


def summarize_csv(paths, kernels: str = "") -> str:
    per = defaultdict(lambda: defaultdict(list))
    for p in paths:
        with open(p) as f:
            for r in csv.DictReader(f):
                per[_clean(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = []
    for n, d in sorted(per.items()):
        if kernels and not re.search(kernels, n):
            continue
        out.append(n)
        mean = {c: sum(v) / len(v) for c, v in d.items()}
        for c, v in sorted(d.items()):
            out.append(f"    {c:28s} {mean[c]:18.1f}   (mean of {len(v)} dispatches)")
        for name, num, den, what in _RATIOS:
            if num in mean and den in mean and mean[den] > 0:
                out.append(f"    = {name:26s} {mean[num] / mean[den]:18.4f}   ({what})")
    return "\n".join(out)


def cmd_summarize(a) -> int:
    paths = [Path(p) for p in a.paths]
    dbs = [p for p in paths if p.suffix == ".db"]
    csvs = [p for p in paths if p.suffix == ".csv"]
    for p in dbs:
        print(summarize_db(p, a.tail_of, a.busy))
        if a.markers:
            print()
            print(summarize_markers(p))
    if csvs:
        print(summarize_csv(csvs, a.kernels))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--workload", default="dam_break",
                   choices=["dam_break", "snow_spheres", "elastic_sphere"])
    r.add_argument("--steps", type=int, default=20)
    r.add_argument("--warmup", type=int, default=5)
    r.add_argument("--option", action="append", help="engine option name=value")
    m = sub.add_parser("mesh")
    m.add_argument("--workload", default="dam_break", choices=["dam_break", "snow_spheres", "elastic_sphere"])
    m.add_argument("--warmup", type=int, default=200)
    m.add_argument("--side", type=int, default=250)
    m.add_argument("--radius", type=int, default=5)
    m.add_argument("--reps", type=int, default=3)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=512)
    c.add_argument("--reps", type=int, default=3)
    s = sub.add_parser("summarize")
    s.add_argument("paths", nargs="+")
    s.add_argument("--tail-of", default="", help="KERNEL:N -- only launches from the N-th last KERNEL on")
    s.add_argument("--kernels", default="", help="regex: counter tables of the matching kernels only")
    s.add_argument("--markers", action="store_true", help="also the roctx ranges of a --marker-trace run")
    s.add_argument("--busy", action="store_true", help="also the time no kernel ran between the launches kept")
    a = ap.parse_args(argv)
    return {"run": cmd_run, "mesh": cmd_mesh, "copy": cmd_copy, "summarize": cmd_summarize}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
