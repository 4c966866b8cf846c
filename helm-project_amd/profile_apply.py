#!/usr/bin/env python3
"""Apply-only driver for profiling the batched apply on one GPU.

Builds the first-order conservative weights of a headline pair (0.25-degree
lat-lon -> C384, or 0.1-degree -> C768), then times Y = A X on k device fields
for the requested numerics modes and prints one JSON line per mode (launch
shape, time per apply from HIP events on the launch stream, algorithmic bytes
and GB/s).  --reps applies follow each timing, so a profiler run
(rocprofv3 --kernel-trace --stats / --pmc) sees a known number of launches
of each mode's shape.

    python3 helm-project_amd/profile_apply.py --pair c384 --fields 64 --modes comp32,acc64,plain
"""
import argparse
import importlib.util
import json
import pathlib
import sys

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
    ap.add_argument("--pair", default="c384", choices=["c384", "c768"])
    ap.add_argument("--fields", type=int, default=64)
    ap.add_argument("--modes", default="comp32,acc64,plain")
    ap.add_argument("--src", default="f32", choices=["f32", "bf16"], help="field storage of the mixed modes")
    ap.add_argument("--dst", default="f32", choices=["f32", "f64"], help="output storage of the mixed modes")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--no-tune", action="store_true", help="keep the library's default launch shapes")
    ap.add_argument("--shape", default=None,
                    help="pin the launch shape 'fields_per_step,field_chunks,staged' (staged: 0 direct, 1 workgroup tiles, 2 wave tiles)")
    ap.add_argument("--sweep", action="store_true", help="time every launch shape and print one line each")
    ap.add_argument("--calibrate", action="store_true",
                    help="end with three copies of 256 Mi FP32 values (1 GiB read, 1 GiB written): the known-size kernel "
                         "against which memory counters are scaled")
    args = ap.parse_args()

    import torch

    _package()
    from helm_project_amd import axis, grids

    axis.init()
    dev = torch.device("cuda:0")
    src_m, dst_m = (grids.latlon_deg(0.25), grids.cubed_sphere(384)) if args.pair == "c384" else (grids.latlon_deg(0.1), grids.cubed_sphere(768))
    S, D = axis.Mesh(src_m), axis.Mesh(dst_m)
    m = axis.generate_weights(S, D, axis.CONSERVATIVE)
    n_src, n_dst, nnz = m.info()
    k = args.fields
    stream = torch.cuda.current_stream()
    gen = torch.Generator(device=dev).manual_seed(1003)
    X64 = torch.randn((k, n_src), generator=gen, device=dev, dtype=torch.float64)
    codes = {"f32": (axis.F32, torch.float32, 4), "bf16": (axis.BF16, torch.bfloat16, 2), "f64": (axis.F64, torch.float64, 8)}
    for name in args.modes.split(","):
        if name == "plain":
            xt, yt, mode = "f64", "f64", axis.PLAIN
        else:
            xt, yt, mode = args.src, args.dst, axis.MIXED if name == "comp32" else axis.MIXED_ACC64
        X = X64.to(codes[xt][1]).contiguous()
        Y = torch.empty((k, n_dst), dtype=codes[yt][1], device=dev)
        nbytes = n_src * k * codes[xt][2] + n_dst * k * codes[yt][2] + nnz * 12 + (n_dst + 1) * 8

        def timed(shape):
            m.batch_apply(X, Y, k, codes[xt][0], codes[yt][0], mode, stream=stream.cuda_stream)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(args.reps):
                m.batch_apply(X, Y, k, codes[xt][0], codes[yt][0], mode, stream=stream.cuda_stream, sync=False)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1) / args.reps
            print(json.dumps({"pair": args.pair, "mode": name, "src": xt, "dst": yt, "fields": k, "shape": shape, "ms": ms,
                              "algorithmic_bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9, "nnz": nnz, "n_src": n_src,
                              "n_dst": n_dst}), flush=True)

        if args.sweep:
            for staged in (1, 2):
                for fu in (1, 2, 4):
                    for chunks in (1, 2, 4, 8):
                        m.set_shape(k, codes[xt][0], codes[yt][0], mode, fu, chunks, staged)
                        timed([fu, chunks, staged])
        elif args.shape:
            shape = [int(v) for v in args.shape.split(",")]
            m.set_shape(k, codes[xt][0], codes[yt][0], mode, *shape)
            timed(shape)
        else:
            shape = None if args.no_tune else list(m.autotune(X, Y, k, codes[xt][0], codes[yt][0], mode, stream=stream.cuda_stream))
            timed(shape)
        del X, Y
    if args.calibrate:
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes_read": a.numel() * 4, "calibration_copy_bytes_written": a.numel() * 4}), flush=True)


if __name__ == "__main__":
    main()
