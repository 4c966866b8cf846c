#!/usr/bin/env python3
"""Kernel-trace summary of a profiled driver run (rocprofv3 --kernel-trace
--output-format csv): the trace is split into clusters wherever the GPU
idles longer than --gap-ms, and for the last --last clusters it prints each
cluster's kernel count, span (first start to last end) and busy time (the
union of the kernel intervals), then one table of every kernel over those
clusters (calls, average and total time, share of the busy time).  With
--walls, the driver's JSON line (its "ms" list: the wall time of each timed
call) is read from that file and each cluster's busy time is given as a
share of the matching call's wall time.

    rocprofv3 --kernel-trace --output-format csv -d OUT -o kt -- python3 helm-project_amd/profile_generate.py --reps 5 --gap-ms 50 > OUT/driver.log
    python3 helm-project_amd/profile_kernels.py OUT --last 5 --walls OUT/driver.log
"""
import argparse
import collections
import csv
import json
import pathlib
import re


def kernels(root):
    out = []
    for f in sorted(pathlib.Path(root).rglob("*kernel_trace.csv")):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                out.append((int(row["Start_Timestamp"]), int(row["End_Timestamp"]), row["Kernel_Name"]))
    out.sort()
    return out


def clusters(ks, gap_ns):
    groups, cur, end = [], [], 0
    for k in ks:
        if cur and k[0] - end > gap_ns:
            groups.append(cur)
            cur = []
        end = max(end, k[1]) if cur else k[1]
        cur.append(k)
    if cur:
        groups.append(cur)
    return groups


def busy_ns(group):
    total, s0, e0 = 0, None, None
    for s, e, _ in group:
        if e0 is None or s > e0:
            if e0 is not None:
                total += e0 - s0
            s0, e0 = s, e
        else:
            e0 = max(e0, e)
    return total + (e0 - s0 if e0 is not None else 0)


def short(name, width=100):
    n = re.sub(r"\(anonymous namespace\)::", "", name)
    n = n.replace("axis_hip::", "").replace("detail::", "").replace("apply_detail::", "")
    i = n.find("(")
    n = n[:i] if i > 0 else n
    return n if len(n) <= width else n[: width - 3] + "..."


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("root", help="rocprofv3 output directory (-d)")
    ap.add_argument("--gap-ms", type=float, default=20.0, help="idle time that separates clusters")
    ap.add_argument("--last", type=int, default=0, help="summarise the last N clusters (0: all)")
    ap.add_argument("--walls", default=None, help="file holding the driver's JSON line with its per-call wall times (ms)")
    ap.add_argument("--top", type=int, default=40)
    args = ap.parse_args()
    groups = clusters(kernels(args.root), int(args.gap_ms * 1e6))
    sel = groups[-args.last:] if args.last > 0 else groups
    walls = None
    if args.walls:
        for line in open(args.walls):
            line = line.strip()
            if line.startswith("{") and '"ms"' in line:
                walls = json.loads(line)["ms"]
    print(f"clusters in the trace: {len(groups)}; summarised: {len(sel)}")
    print()
    print("| cluster | kernels | span ms | busy ms | wall ms | busy / wall |")
    print("|---|---|---|---|---|---|")
    for i, g in enumerate(sel):
        span = (max(e for _, e, _ in g) - g[0][0]) / 1e6
        busy = busy_ns(g) / 1e6
        w = walls[len(walls) - len(sel) + i] if walls and len(walls) >= len(sel) else None
        ws = f"{w:.3f}" if w is not None else "-"
        fr = f"{busy / w:.2f}" if w else "-"
        print(f"| {i} | {len(g)} | {span:.3f} | {busy:.3f} | {ws} | {fr} |")
    agg = collections.defaultdict(lambda: [0, 0])
    for g in sel:
        for s, e, n in g:
            agg[short(n)][0] += 1
            agg[short(n)][1] += e - s
    tot = sum(v[1] for v in agg.values())
    print()
    print(f"kernel time per cluster: {tot / 1e6 / max(len(sel), 1):.3f} ms")
    print()
    print("| kernel | calls per cluster | avg us | ms per cluster | % |")
    print("|---|---|---|---|---|")
    nc = max(len(sel), 1)
    for n, (c, t) in sorted(agg.items(), key=lambda x: -x[1][1])[: args.top]:
        print(f"| `{n}` | {c / nc:.1f} | {t / c / 1e3:.1f} | {t / 1e6 / nc:.3f} | {100.0 * t / tot:.1f} |")


if __name__ == "__main__":
    main()
