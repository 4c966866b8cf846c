"""Trains the candidate-proposer network for spherical Voronoi sources and a
C384 destination and writes helm-project_amd/data/proposer_voronoi8_c384.npz.

Training set: 20 jittered level-8 icosahedral Voronoi meshes (655,362 cells,
seeds 5000-5019), 100,000 sampled C384 cells each (2,000,000 pairs), labels
from axis.proposer_training_data.  Training runs on the GPU through the
library's own forward / backward / Adam kernels.  The held-out mesh (seed
5020) is then regridded with and without the proposer.

    python helm-project_amd/train_proposer.py [--level 8] [--meshes 20] [--per-mesh 100000]
"""
import argparse
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--meshes", type=int, default=20)
    ap.add_argument("--per-mesh", type=int, default=100_000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--dst", type=int, default=384)
    ap.add_argument("--out", default=str(ROOT / "helm-project_amd" / "data" / "proposer_voronoi8_c384.npz"))
    args = ap.parse_args()

    import __graft_entry__ as ge

    ge._import_package()
    from helm_project_amd import axis, grids, proposer

    # meshes first, on the CPU, before anything touches the GPU
    t = time.perf_counter()
    seeds = list(range(5000, 5000 + args.meshes))
    meshes = proposer.voronoi_meshes(args.level, seeds, workers=8)
    log(f"{len(meshes)} Voronoi level-{args.level} meshes in {time.perf_counter() - t:.1f} s")
    dst = grids.cubed_sphere(args.dst)
    n_dst = grids.n_cells(dst)
    scale = proposer.feature_scale(n_dst)

    import torch

    axis.init()
    t = time.perf_counter()
    X, T = proposer.training_set(meshes, dst, scale, per_mesh=min(args.per_mesh, n_dst))
    torch.cuda.synchronize()
    log(f"training set {tuple(X.shape)} in {time.perf_counter() - t:.1f} s")
    del meshes
    t = time.perf_counter()
    net, losses = proposer.train(X, T, scale, epochs=args.epochs, batch=8192, lr=2e-3, log=log)
    torch.cuda.synchronize()
    log(f"trained in {time.perf_counter() - t:.1f} s")
    proposer.save(net, args.out, level=args.level, seeds=seeds, per_mesh=args.per_mesh, epochs=args.epochs,
                  losses=losses, dst=args.dst, tau=proposer.TAU)
    log("wrote", args.out)

    # held-out mesh
    src = grids.jittered_voronoi(args.level, 5000 + args.meshes)
    S, D = axis.Mesh(src), axis.Mesh(dst)
    for rep in range(2):
        t = time.perf_counter()
        plain = axis.generate_weights(S, D, axis.CONSERVATIVE)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t
        t = time.perf_counter()
        prop, st = axis.generate_weights_proposed(S, D, net)
        torch.cuda.synchronize()
        t_prop = time.perf_counter() - t
        ra, ca, va = plain.csr()
        rb, cb, vb = prop.csr()
        same = bool((ra == rb).all() and (ca == cb).all() and (va == vb).all())
        log(f"held-out: plain {t_plain * 1e3:.1f} ms ({plain.stats()['candidates']} pairs), proposed {t_prop * 1e3:.1f} ms {st}, "
            f"identical={same}")
        plain.close()
        prop.close()


if __name__ == "__main__":
    main()
