"""Summaries of rocprofv3 output for the profiles/ folder.

  python -m nenya_amd.profile_summary stats KERNEL_STATS_CSV [--top N] [--cmd CMD]
      Markdown table of a `rocprofv3 --kernel-trace --stats --output-format csv`
      kernel_stats.csv (per-kernel calls, total and average duration).

  python -m nenya_amd.profile_summary traffic --fetch F.csv --write W.csv
          --kernel NAME --key KEY --calib-bytes B --out profiles/traffic.json
      HBM bytes per launch of kernel NAME from two `rocprofv3 --pmc FETCH_SIZE`
      / `--pmc WRITE_SIZE` counter_collection.csv files (KiB per dispatch).
      Each counter is scaled by what it reads for the calibration copy of B
      bytes in the same run (the dispatch of the calibration kernel, k_copy_out
      by default, with the largest WRITE_SIZE), so the factor absorbs the
      counter's units.  The result is merged into the JSON file under KEY.

  python -m nenya_amd.profile_summary kernels --stats KERNEL_STATS_CSV
          [--mfma M.csv] [--fetch F.csv --write W.csv] [--calib-bytes B]
          [--kernels K1,K2,...] [--cmd CMD]
      Markdown table per kernel: calls, average duration (kernel trace), and
      from counter runs the int8 MFMA operations per dispatch
      (SQ_INSTS_VALU_MFMA_MOPS_I8 x 512) as TOP/s against the 5.0 POP/s dense
      int8 peak, the MFMA busy share (SQ_VALU_MFMA_BUSY_CYCLES over
      GRBM_GUI_ACTIVE x 1024 SIMDs), and HBM bytes per dispatch as GB/s
      against 8 TB/s.
"""
import argparse
import csv
import json
import os
import sys


def short(name):
    n = name.replace("void ", "").replace("nenya::", "")
    return n.split("(")[0]


def stats(args):
    rows = list(csv.DictReader(open(args.csv)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = []
    if args.cmd:
        out.append(f"Command: `{args.cmd}`\n")
    out.append("| kernel | calls | total ms | avg us | min us | max us | % |")
    out.append("|---|---:|---:|---:|---:|---:|---:|")
    for r in rows[: args.top]:
        out.append(f"| {short(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{100 * float(r['TotalDurationNs']) / total:.1f} |")
    print("\n".join(out))


def load_counters(path):
    return list(csv.DictReader(open(path)))


def calib_scales(fr, wr, calib_kernel, calib_bytes):
    """FETCH_SIZE / WRITE_SIZE scale factors (bytes per counted unit... as KiB
    ratios) from the calibration copy: the dispatch of `calib_kernel` with
    the largest WRITE_SIZE."""
    cal_w = max((r for r in wr if calib_kernel in r["Kernel_Name"]), key=lambda r: float(r["Counter_Value"]))
    cal_f = max((r for r in fr if calib_kernel in r["Kernel_Name"]), key=lambda r: float(r["Counter_Value"]))
    kib = calib_bytes / 1024.0
    return kib / float(cal_f["Counter_Value"]), kib / float(cal_w["Counter_Value"]), cal_f, cal_w


def _per_kernel(rows, counter=None):
    out = {}
    for r in rows:
        if counter and r["Counter_Name"] != counter:
            continue
        out.setdefault(short(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return out


def kernels(args):
    st = {short(r["Name"]): r for r in csv.DictReader(open(args.stats))}
    names = args.kernels.split(",") if args.kernels else sorted(st, key=lambda n: -float(st[n]["TotalDurationNs"]))
    mf = load_counters(args.mfma) if args.mfma else []
    mops, busy, act = (_per_kernel(mf, c) for c in ("SQ_INSTS_VALU_MFMA_MOPS_I8", "SQ_VALU_MFMA_BUSY_CYCLES",
                                                   "GRBM_GUI_ACTIVE"))
    fet = wri = {}
    fs = ws = None
    if args.fetch and args.write:
        fr, wr = load_counters(args.fetch), load_counters(args.write)
        fs, ws, cf, cw = calib_scales(fr, wr, args.calib_kernel, args.calib_bytes)
        fet, wri = _per_kernel(fr), _per_kernel(wr)
    out = []
    if args.cmd:
        out.append(f"Command: `{args.cmd}`\n")
    if fs:
        out.append(f"Traffic scale (calibration copy of {args.calib_bytes} bytes by {args.calib_kernel}): "
                   f"FETCH_SIZE x{fs:.3f}, WRITE_SIZE x{ws:.3f}.\n")
    out.append("| kernel | calls | avg us | int8 TOP/s | of 5.0 POP/s | MFMA busy | HBM MB/launch | GB/s | of 8 TB/s |")
    out.append("|---|---:|---:|---:|---:|---:|---:|---:|---:|")
    mean = lambda v: sum(v) / len(v)  # noqa: E731
    for n in names:
        matches = [k for k in st if k == n or k.startswith(n + "<")]
        for k in matches:
            r = st[k]
            us = float(r["AverageNs"]) / 1e3
            cells = [k, r["Calls"], f"{us:.1f}"]
            if k in mops and mean(mops[k]) > 0:
                ops = mean(mops[k]) * 512
                tops = ops / (us * 1e-6) / 1e12
                cells += [f"{tops:.0f}", f"{tops / 5000:.3f}",
                          f"{100 * mean(busy[k]) / (mean(act[k]) * 1024):.1f}%" if k in busy and k in act else "-"]
            else:
                cells += ["-", "-", "-"]
            if k in fet and k in wri:
                b = (mean(fet[k]) * fs + mean(wri[k]) * ws) * 1024
                gbs = b / (us * 1e-6) / 1e9
                cells += [f"{b / 1e6:.2f}", f"{gbs:.0f}", f"{gbs / 8000:.3f}"]
            else:
                cells += ["-", "-", "-"]
            out.append("| " + " | ".join(cells) + " |")
    print("\n".join(out))


def traffic(args):
    fr, wr = load_counters(args.fetch), load_counters(args.write)
    f_scale, w_scale, cal_f, cal_w = calib_scales(fr, wr, args.calib_kernel, args.calib_bytes)
    cal_id = cal_w["Dispatch_Id"]
    # a comma-separated list averages over the dispatches of several kernels
    # (the split and persistent forms of one operation)
    names = args.kernel.split(",")

    def hit(r):
        n = short(r["Kernel_Name"])
        return any(n == k or n.startswith(k + "<") for k in names)

    fk = [float(r["Counter_Value"]) for r in fr if hit(r)]
    wk = [float(r["Counter_Value"]) for r in wr if hit(r)]
    if not fk or len(fk) != len(wk):
        sys.exit(f"kernel {args.kernel}: {len(fk)} FETCH and {len(wk)} WRITE dispatches")
    fetch = sum(fk) / len(fk) * f_scale * 1024
    write = sum(wk) / len(wk) * w_scale * 1024
    entry = {"kernel": args.kernel, "dispatches": len(fk), "fetch_bytes_per_launch": round(fetch),
             "write_bytes_per_launch": round(write), "bytes_per_launch": round(fetch + write),
             "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
             "calibration": f"copyBuffer dispatch {cal_id} of {args.calib_bytes} bytes read "
                            f"{cal_f['Counter_Value']} KiB FETCH_SIZE, {cal_w['Counter_Value']} KiB WRITE_SIZE"}
    d = {}
    if os.path.exists(args.out):
        d = json.load(open(args.out))
    d[args.key] = entry
    with open(args.out, "w") as fh:
        json.dump(d, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print(json.dumps(entry))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd_name", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=40)
    s.add_argument("--cmd", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--key", required=True)
    t.add_argument("--calib-bytes", type=int, default=1 << 28)
    t.add_argument("--out", required=True)
    t.add_argument("--calib-kernel", default="k_copy_out")
    k = sub.add_parser("kernels")
    k.add_argument("--stats", required=True)
    k.add_argument("--mfma")
    k.add_argument("--fetch")
    k.add_argument("--write")
    k.add_argument("--calib-bytes", type=int, default=1 << 28)
    k.add_argument("--calib-kernel", default="k_copy_out")
    k.add_argument("--kernels", default="")
    k.add_argument("--cmd", default="")
    args = ap.parse_args()
    {"stats": stats, "traffic": traffic, "kernels": kernels}[args.cmd_name](args)


if __name__ == "__main__":
    main()
