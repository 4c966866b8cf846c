#!/usr/bin/env python3
"""FP64 operation count of the clip pass of a conservative generation, from
the oracle's instrumented sequential loop (every pair that passes the cap
filter is clipped and its area taken; +, -, *, /, sqrt and atan2 count one
each, in the GPU kernel's formulation).  Prints one line per key, in the form
the benchmark reads (`clip_fp64_ops_per_generation = N`):

    python3 oracle/count_clip_ops.py --pair c384
"""
import argparse
import importlib.util
import pathlib
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
sys.path.insert(0, str(HERE))


def _package():
    spec = importlib.util.spec_from_file_location("helm_project_amd", ROOT / "helm-project_amd" / "__init__.py",
                                                  submodule_search_locations=[str(ROOT / "helm-project_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["helm_project_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pair", default="c384", choices=["c24", "c384", "c768"])
    args = ap.parse_args()
    _package()
    from helm_project_amd import grids
    import oracle_py as O

    src, dst = {"c24": (grids.latlon_deg(2.0), grids.cubed_sphere(24)), "c384": (grids.latlon_deg(0.25), grids.cubed_sphere(384)),
                "c768": (grids.latlon_deg(0.1), grids.cubed_sphere(768))}[args.pair]
    t = time.perf_counter()
    r = O.generate(src, dst, 2, count_ops=True)
    ops = r["ops"]
    print(f"pair = {args.pair}")
    print(f"oracle_seconds = {time.perf_counter() - t:.2f}")
    print(f"clipped_pairs = {ops['pairs']}")
    print(f"clip_edges = {ops['edges']}")
    print(f"edge_intersections = {ops['intersections']}")
    print(f"area_fan_terms = {ops['fan_terms']}")
    print(f"nnz = {len(r['val'])}")
    print(f"clip_fp64_ops_per_generation = {int(ops['flops'])}")
    print(f"clip_fp64_ops_per_pair = {ops['flops'] / max(ops['pairs'], 1):.1f}")


if __name__ == "__main__":
    main()
