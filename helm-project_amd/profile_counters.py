#!/usr/bin/env python3
"""Per-kernel averages of a rocprofv3 counter run (--pmc ... --output-format
csv): for every kernel whose name matches --kernel, the dispatch count, grid
and workgroup size, registers, LDS, average duration and the average of each
counter per dispatch; --grid keeps only dispatches of one grid size (one launch
shape of a kernel that several shapes share).

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o run -- python3 helm-project_amd/profile_apply.py ...
    python3 helm-project_amd/profile_counters.py OUT --kernel spmm_tile_kernel
"""
import argparse
import collections
import csv
import pathlib
import re
import statistics


def counter_files(root):
    return sorted(pathlib.Path(root).rglob("*counter_collection.csv"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("root", help="rocprofv3 output directory (-d)")
    ap.add_argument("--kernel", default=".", help="regex on the kernel name")
    ap.add_argument("--grid", type=int, default=None, help="keep only dispatches of this grid size")
    args = ap.parse_args()
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    meta = {}
    for f in counter_files(args.root):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                name = row["Kernel_Name"]
                if not re.search(args.kernel, name):
                    continue
                grid = int(row["Grid_Size"])
                if args.grid is not None and grid != args.grid:
                    continue
                key = (name, row["Dispatch_Id"])
                per[key][row["Counter_Name"]] += float(row["Counter_Value"])
                meta[key] = (grid, int(row["Workgroup_Size"]), int(row.get("VGPR_Count", 0) or 0), int(row.get("LDS_Block_Size", 0) or 0),
                             (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3)
    kernels = collections.defaultdict(list)
    for (name, disp), cs in per.items():
        kernels[name].append((meta[(name, disp)], cs))
    for name, items in sorted(kernels.items()):
        grids = sorted({m[0] for m, _ in items})
        m0 = items[0][0]
        durs = [m[4] for m, _ in items]
        print(f"kernel: {name}")
        print(f"  dispatches {len(items)}  grid sizes {grids}  workgroup {m0[1]}  vgpr {m0[2]}  lds {m0[3]}")
        print(f"  duration_us mean {statistics.mean(durs):.1f} median {statistics.median(durs):.1f}")
        names = sorted({n for _, cs in items for n in cs})
        for n in names:
            vals = [cs.get(n, 0.0) for _, cs in items]
            print(f"  {n:34s} mean {statistics.mean(vals):18.1f}  median {statistics.median(vals):18.1f}")


if __name__ == "__main__":
    main()
