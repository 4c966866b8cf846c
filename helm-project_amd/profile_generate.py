#!/usr/bin/env python3
"""Weight-generation driver for profiling on one GPU.

Builds the meshes of a headline pair once, then runs --reps generations of
the requested method (after --warmup untimed ones) and prints one JSON line
with the wall time of each, so that a kernel/runtime trace of the run covers
a known number of generations.  --gap-ms idles the GPU between generations
(outside the timed calls), so that a kernel trace splits into one cluster
per generation (profile_kernels.py --gap-ms):

    python3 helm-project_amd/profile_generate.py --pair c384 --method conservative --reps 3 --gap-ms 50
"""
import argparse
import importlib.util
import json
import pathlib
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent


def _package():
    if "helm_project_amd" in sys.modules:
        return sys.modules["helm_project_amd"]
    spec = importlib.util.spec_from_file_location("helm_project_amd", HERE / "__init__.py", submodule_search_locations=[str(HERE)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["helm_project_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pair", default="c384", choices=["c384", "c768", "voronoi8"])
    ap.add_argument("--method", default="conservative", choices=["conservative", "bilinear"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--gap-ms", type=float, default=0.0, help="idle time after each generation (not timed)")
    args = ap.parse_args()

    import torch

    _package()
    from helm_project_amd import axis, grids

    axis.init()
    if args.pair == "c384":
        src, dst = grids.latlon_deg(0.25), grids.cubed_sphere(384)
    elif args.pair == "c768":
        src, dst = grids.latlon_deg(0.1), grids.cubed_sphere(768)
    else:
        src, dst = grids.jittered_voronoi(8, 5020), grids.cubed_sphere(384)
    S, D = axis.Mesh(src), axis.Mesh(dst)
    method = axis.CONSERVATIVE if args.method == "conservative" else axis.BILINEAR
    for _ in range(args.warmup):
        axis.generate_weights(S, D, method).close()
    torch.cuda.synchronize()
    if args.gap_ms > 0:
        time.sleep(args.gap_ms * 1e-3)
    times, phases = [], []
    for _ in range(args.reps):
        t = time.perf_counter()
        m = axis.generate_weights(S, D, method)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t) * 1e3)
        nnz = m.info()[2]
        phases.append(m.phase_ms())
        m.close()
        if args.gap_ms > 0:
            torch.cuda.synchronize()
            time.sleep(args.gap_ms * 1e-3)
    print(json.dumps({"pair": args.pair, "method": args.method, "ms": times, "nnz": nnz, "phase_gpu_ms": phases}), flush=True)


if __name__ == "__main__":
    main()
