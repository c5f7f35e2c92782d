"""Summarise rocprofv3 CSV output into profiles/ (text only).

Usage (after a GPU run that wrote the rocprofv3 directories):
  python compute-vp9_amd/profsum.py --name NAME --workload TEXT --command TEXT \
      --stats DIR/*_kernel_stats.csv --fetch F.csv --write W.csv \
      --copy-fetch CF.csv --copy-write CW.csv [--copy-bytes N]

FETCH_SIZE / WRITE_SIZE are reported by the hardware in KiB per dispatch.
Each is scaled by the factor that makes the calibration copy kernel (which
reads and writes exactly --copy-bytes per dispatch) read its known size.
Writes profiles/NAME.md and merges the per-kernel bytes into
profiles/traffic.json, keyed by workload.
"""
import argparse
import collections
import csv
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(ROOT, "profiles")


def short(name):
    n = name.split("(")[0]
    return n.replace("void ", "")


def per_dispatch(path):
    agg = collections.defaultdict(lambda: [0.0, 0])
    for r in csv.DictReader(open(path)):
        a = agg[short(r["Kernel_Name"])]
        a[0] += float(r["Counter_Value"])
        a[1] += 1
    return {k: (v[0] / v[1], v[1]) for k, v in agg.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--name", required=True)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--command", required=True, action="append")
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--note", action="append", default=[])
    a = ap.parse_args()

    rows = list(csv.DictReader(open(a.stats)))
    fetch = per_dispatch(a.fetch) if a.fetch else {}
    write = per_dispatch(a.write) if a.write else {}
    ff = wf = None
    if a.copy_fetch and a.copy_write:
        cf = per_dispatch(a.copy_fetch)["k_copy_probe"][0]
        cw = per_dispatch(a.copy_write)["k_copy_probe"][0]
        ff = (a.copy_bytes / 1024) / cf
        wf = (a.copy_bytes / 1024) / cw

    lines = [f"# {a.name}", "", f"Workload: {a.workload}", "", "Commands (on a machine with the GPU, "
             "after `cd /tmp && export TMPDIR=/tmp`; `$R` is the repository root):", ""]
    for c in a.command:
        lines += ["```", c, "```"]
    lines += ["", "## Kernel durations (rocprofv3 --kernel-trace --stats)", "",
              "| kernel | calls | avg us | min us | max us | share % |", "|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| {short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                     f"{float(r['Percentage']):.2f} |")
    traffic = {}
    if fetch and write and ff:
        lines += ["", "## HBM-side traffic per dispatch (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)", "",
                  f"Calibration: the copy kernel moves {a.copy_bytes} B in and out per dispatch; "
                  f"FETCH_SIZE scale {ff:.4f}, WRITE_SIZE scale {wf:.4f}.", "",
                  "| kernel | dispatches | FETCH KiB (raw) | WRITE KiB (raw) | fetch MB | write MB | traffic MB |",
                  "|---|---|---|---|---|---|---|"]
        for k in sorted(fetch, key=lambda k: -fetch[k][0]):
            if k not in write or k.startswith("__amd"):
                continue
            fb = fetch[k][0] * 1024 * ff
            wb = write[k][0] * 1024 * wf
            traffic[k] = {"fetch_bytes": round(fb), "write_bytes": round(wb),
                          "traffic_bytes": round(fb + wb)}
            lines.append(f"| {k} | {fetch[k][1]} | {fetch[k][0]:.1f} | {write[k][0]:.1f} | "
                         f"{fb / 1e6:.3f} | {wb / 1e6:.3f} | {(fb + wb) / 1e6:.3f} |")
    for n in a.note:
        lines += ["", n]
    os.makedirs(PROFILES, exist_ok=True)
    with open(os.path.join(PROFILES, a.name + ".md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    if traffic:
        path = os.path.join(PROFILES, "traffic.json")
        data = json.load(open(path)) if os.path.exists(path) else {}
        data[a.workload] = {"kernels": traffic, "fetch_scale": ff, "write_scale": wf,
                            "profile": a.name + ".md"}
        with open(path, "w") as f:
            json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
