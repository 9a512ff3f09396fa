"""Summaries of rocprofv3 output for the profiles/ folder.

  python -m nenya_amd.profile_summary stats KERNEL_STATS_CSV [--top N] [--cmd CMD]
      Markdown table of a `rocprofv3 --kernel-trace --stats --output-format csv`
      kernel_stats.csv (per-kernel calls, total and average duration).

  python -m nenya_amd.profile_summary traffic --fetch F.csv --write W.csv
          --kernel NAME --key KEY --calib-bytes B --out profiles/traffic.json
      HBM bytes per launch of kernel NAME from two `rocprofv3 --pmc FETCH_SIZE`
      / `--pmc WRITE_SIZE` counter_collection.csv files (KiB per dispatch).
      Each counter is scaled by what it reads for the calibration copy of B
      bytes in the same run (the copyBuffer dispatch with the largest
      WRITE_SIZE), so the factor absorbs the counter's units.  The result is
      merged into the JSON file under KEY.
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


def traffic(args):
    fr, wr = load_counters(args.fetch), load_counters(args.write)
    cal_w = max((r for r in wr if "copyBuffer" in r["Kernel_Name"]), key=lambda r: float(r["Counter_Value"]))
    cal_id = cal_w["Dispatch_Id"]
    cal_f = next(r for r in fr if r["Dispatch_Id"] == cal_id and "copyBuffer" in r["Kernel_Name"])
    kib = args.calib_bytes / 1024.0
    f_scale = kib / float(cal_f["Counter_Value"])
    w_scale = kib / float(cal_w["Counter_Value"])
    fk = [float(r["Counter_Value"]) for r in fr if short(r["Kernel_Name"]).startswith(args.kernel + "<")
          or short(r["Kernel_Name"]) == args.kernel]
    wk = [float(r["Counter_Value"]) for r in wr if short(r["Kernel_Name"]).startswith(args.kernel + "<")
          or short(r["Kernel_Name"]) == args.kernel]
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
    args = ap.parse_args()
    stats(args) if args.cmd_name == "stats" else traffic(args)


if __name__ == "__main__":
    main()
