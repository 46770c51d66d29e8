"""Short text summaries of rocprofv3 CSV output (``--output-format csv``).

  python -m fv3atm_amd.rocprof_summary stats DIR
      per-kernel calls, total and mean duration from the kernel trace
  python -m fv3atm_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR --kernel K
          --launches L --calib-bytes B [--tag "k=v ..."]
      HBM bytes of every dispatch whose name contains K (FETCH_SIZE run plus
      WRITE_SIZE run, each calibrated as below), summed and divided by L
      launches; prints a 'bench-traffic:' line carrying the tag
  python -m fv3atm_amd.rocprof_summary spans DIR --kernel K --group G
      wall span (first start to last end) of every G consecutive dispatches
      whose name contains K, ordered by start: one launch made of G kernels
      that may run concurrently (mean, min, max)
  python -m fv3atm_amd.rocprof_summary valu SQ_DIR TRACE_DIR --kernel K --launches L
          [--only S] [--tag "k=v ..."]
      per-kernel VALU issue rate: the mean SQ_INSTS_VALU per dispatch of the
      counter run SQ_DIR over the mean duration per dispatch of the kernel
      trace TRACE_DIR (a run whose kernels do not overlap), against the MI355X
      issue peak (256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64
      instruction); then the VALU instructions of every dispatch whose name
      contains K summed and divided by L launches, as a 'bench-valu:' line
  python -m fv3atm_amd.rocprof_summary counters DIR [--calib-bytes B]
      per-kernel mean of every collected counter per dispatch; with
      --calib-bytes, the counter of the calibration kernel (the one whose
      name contains --calib-name, default "bitwise_not") is taken to stand
      for B bytes, and every other kernel's counter is scaled by the same
      factor into HBM bytes per dispatch.

The summaries print the figures only; the rocprofv3 command that made DIR
is recorded by the caller beside them.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re
import sys
from collections import defaultdict


def _rows(d, suffix):
    out = []
    for p in sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)):
        with open(p, newline="") as f:
            out.extend(csv.DictReader(f))
    if not out:
        raise SystemExit(f"no *{suffix} under {d}")
    return out


def _col(row, *names):
    low = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in low:
            return row[low[n.lower()]]
    raise KeyError(f"none of {names} in {list(row)}")


def short(name):
    """Kernel name without its argument list, keeping template arguments."""
    s = re.sub(r"^void\s+", "", name.strip()).replace("(anonymous namespace)", "{anon}")
    depth, cut = 0, len(s)
    for i, ch in enumerate(s):  # first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return s[:cut]


def clip(name, n=96):
    return name if len(name) <= n else name[: n - 3] + "..."


def kernel_stats(d):
    agg = defaultdict(lambda: [0, 0.0])
    for r in _rows(d, "kernel_trace.csv"):
        n = short(_col(r, "Kernel_Name"))
        t = float(_col(r, "End_Timestamp")) - float(_col(r, "Start_Timestamp"))
        agg[n][0] += 1
        agg[n][1] += t
    return agg


def spans(d, kernel, group):
    """Wall spans (ns) of consecutive groups of `group` dispatches of kernels
    whose short name contains `kernel`, in start order."""
    ev = sorted((float(_col(r, "Start_Timestamp")), float(_col(r, "End_Timestamp")))
                for r in _rows(d, "kernel_trace.csv") if kernel in short(_col(r, "Kernel_Name")))
    out = []
    for i in range(0, len(ev) - group + 1, group):
        g = ev[i:i + group]
        out.append(max(e for _, e in g) - min(b for b, _ in g))
    return out


def counter_means(d):
    """{(kernel, counter): [dispatches, sum]} over dispatches (a counter that a
    dispatch reports once per agent/dimension is summed within the dispatch)."""
    per = defaultdict(float)
    names = {}
    for r in _rows(d, "counter_collection.csv"):
        key = (_col(r, "Dispatch_Id"), _col(r, "Counter_Name"))
        per[key] += float(_col(r, "Counter_Value"))
        names[key[0]] = short(_col(r, "Kernel_Name"))
    agg = defaultdict(lambda: [0, 0.0])
    for (disp, cn), v in per.items():
        a = agg[(names[disp], cn)]
        a[0] += 1
        a[1] += v
    return agg


def calibrated_sum(d, counter, kernel, calib_name, calib_bytes):
    """(sum over dispatches of kernels named like `kernel`, dispatch count,
    calibration factor) of one counter, in bytes."""
    agg = counter_means(d)
    factor = None
    for (n, cn), (c, s) in agg.items():
        if cn == counter and calib_name.lower() in n.lower():
            factor = calib_bytes / (s / c * 1024.0)
    if factor is None:
        raise SystemExit(f"no calibration kernel '{calib_name}' with {counter} under {d}")
    tot, cnt = 0.0, 0
    for (n, cn), (c, s) in agg.items():
        if cn == counter and kernel in n:
            tot += s * 1024.0 * factor
            cnt += c
    return tot, cnt, factor


VALU_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second, MI355X


def valu_table(sq_dir, trace_dir):
    """[(kernel, mean ms, VALU, SALU, VALU G/s, fraction of peak)] for every
    kernel with both a SQ_INSTS_VALU count and a trace duration."""
    agg = counter_means(sq_dir)
    st = kernel_stats(trace_dir)
    out = []
    for (n, cn), (c, v) in sorted(agg.items()):
        if cn != "SQ_INSTS_VALU" or n not in st:
            continue
        calls, tot = st[n]
        dur = tot / calls * 1e-9
        valu = v / c
        sc, sv = agg.get((n, "SQ_INSTS_SALU"), (1, 0.0))
        out.append((n, dur * 1e3, valu, sv / sc, valu / dur / 1e9, valu / dur / VALU_PEAK))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "counters", "traffic", "spans", "valu"])
    ap.add_argument("dir")
    ap.add_argument("dir2", nargs="?")
    ap.add_argument("--kernel", default="k_fused_tp")
    ap.add_argument("--launches", type=int, default=1)
    ap.add_argument("--tag", default="")
    ap.add_argument("--group", type=int, default=1)
    ap.add_argument("--calib-bytes", type=float, default=0.0)
    ap.add_argument("--calib-name", default="bitwise_not")
    ap.add_argument("--top", type=int, default=12)
    ap.add_argument("--only", default="", help="valu: list only kernels whose name contains this")
    a = ap.parse_args(argv)
    w = sys.stdout.write
    if a.mode == "stats":
        agg = kernel_stats(a.dir)
        tot = sum(v[1] for v in agg.values())
        w("| kernel | calls | total ms | mean us | share |\n|---|---:|---:|---:|---:|\n")
        for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[: a.top]:
            w(f"| `{clip(n)}` | {c} | {t / 1e6:.3f} | {t / c / 1e3:.1f} | {100 * t / tot:.1f}% |\n")
        return
    if a.mode == "spans":
        sp = spans(a.dir, a.kernel, a.group)
        if not sp:
            raise SystemExit(f"no dispatches of {a.kernel}")
        w(f"{len(sp)} launches of {a.group} `{a.kernel}` kernels: span mean {sum(sp) / len(sp) / 1e6:.3f} ms, "
          f"min {min(sp) / 1e6:.3f} ms, max {max(sp) / 1e6:.3f} ms\n")
        return
    if a.mode == "valu":
        rows = valu_table(a.dir, a.dir2)
        w("| kernel | mean ms | VALU | SALU | VALU G/s | VALU busy |\n|---|---:|---:|---:|---:|---:|\n")
        for n, ms, v, sv, g, f in rows:
            if a.only in n:
                w(f"| `{clip(n)}` | {ms:.3f} | {v:.4e} | {sv:.4e} | {g:.0f} | {100 * f:.0f}% |\n")
        agg = counter_means(a.dir)
        tot = sum(v for (n, cn), (c, v) in agg.items() if cn == "SQ_INSTS_VALU" and a.kernel in n)
        w(f"VALU instructions of `{a.kernel}` kernels per launch ({a.launches} launches): {tot / a.launches:.4e}\n")
        w(f"bench-valu: {a.tag} valu_per_pass={tot / a.launches:.0f}\n")
        return
    if a.mode == "traffic":
        f, nf, ff = calibrated_sum(a.dir, "FETCH_SIZE", a.kernel, a.calib_name, a.calib_bytes)
        wr, nw, fw = calibrated_sum(a.dir2, "WRITE_SIZE", a.kernel, a.calib_name, a.calib_bytes)
        per = (f + wr) / a.launches
        w(f"kernels matching `{a.kernel}`: {nf} dispatches (FETCH run), {nw} (WRITE run)\n")
        w(f"FETCH_SIZE factor {ff:.3f}: {f:.4e} B; WRITE_SIZE factor {fw:.3f}: {wr:.4e} B\n")
        w(f"HBM bytes per launch ({a.launches} launches): {per:.4e}\n")
        w(f"bench-traffic: {a.tag} bytes_per_pass={per:.0f}\n")
        return
    agg = counter_means(a.dir)
    factor = {}
    if a.calib_bytes > 0:
        for (n, cn), (c, s) in agg.items():
            if a.calib_name.lower() in n.lower():
                factor[cn] = a.calib_bytes / (s / c * 1024.0)  # counters read KiB
                w(f"calibration `{n}` {cn}: {s / c:.0f} KiB per dispatch for {a.calib_bytes:.0f} B "
                  f"-> factor {factor[cn]:.3f}\n")
    w("| kernel | counter | dispatches | mean KiB/dispatch | HBM bytes/dispatch |\n"
      "|---|---|---:|---:|---:|\n")
    for (n, cn), (c, s) in sorted(agg.items()):
        m = s / c
        hb = f"{m * 1024.0 * factor[cn]:.4e}" if cn in factor else "-"
        w(f"| `{clip(n)}` | {cn} | {c} | {m:.1f} | {hb} |\n")


if __name__ == "__main__":
    main()
