"""Training and evaluation helpers of the learned candidate proposer.

The proposer network (axis.MLP, 6 -> 128 -> 128 -> 4) maps a destination
cell's centroid xyz and box extents to a search cap for the conservative
broad phase (axis.generate_weights_proposed).  Its training pairs come from
the library itself (axis.proposer_training_data: the containment-guard
radius of every destination cell against a source mesh), and it is trained
on the GPU by the library's own BF16-MFMA forward / backward and Adam
kernels.
"""
import math
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import axis, grids

TAU = 0.95   # quantile of the guard radius the network is trained to predict
WC = 1.0     # weight of the centre-offset error


def feature_scale(n_dst: int) -> float:
    """Length unit of the features and outputs: the mean cell spacing of the destination mesh (radians)."""
    return math.sqrt(4.0 * math.pi / n_dst)


def _voronoi(args):
    level, seed = args
    return grids.jittered_voronoi(level, seed)


def voronoi_meshes(level: int, seeds, workers: int = 8):
    """Jittered icosahedral Voronoi meshes for several seeds (CPU, in parallel)."""
    seeds = list(seeds)
    if workers <= 1 or len(seeds) == 1:
        return [grids.jittered_voronoi(level, s) for s in seeds]
    with ProcessPoolExecutor(max_workers=min(workers, len(seeds))) as ex:
        return list(ex.map(_voronoi, [(level, s) for s in seeds]))


def training_set(src_meshes, dst_mesh, scale: float, per_mesh: int, seed: int = 0):
    """(X [n, 6], T [n, 4]) device tensors: `per_mesh` sampled destination
    cells of every (source, destination) pair."""
    import torch

    D = axis.Mesh(dst_mesh)
    n_dst = grids.n_cells(dst_mesh)
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    xs, ts = [], []
    for m in src_meshes:
        S = axis.Mesh(m)
        x = torch.empty((n_dst, 6), dtype=torch.float32, device="cuda")
        t = torch.empty((n_dst, 4), dtype=torch.float32, device="cuda")
        axis.proposer_training_data(S, D, scale, x, t)
        idx = torch.randperm(n_dst, generator=g)[:per_mesh].cuda()
        xs.append(x[idx])
        ts.append(t[idx])
        S.close()
    D.close()
    return torch.cat(xs), torch.cat(ts)


def train(X, T, scale: float, epochs: int = 8, batch: int = 8192, lr: float = 2e-3, seed: int = 1, log=None):
    """Adam training of a fresh network on (X, T), shuffled every epoch."""
    import torch

    net = axis.MLP(seed=seed, feature_scale=scale)
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    n = X.shape[0]
    losses = []
    for e in range(epochs):
        perm = torch.randperm(n, generator=g).cuda()
        lr_e = lr * (0.5 ** (e // max(1, epochs // 3)))
        losses.append(net.train(X[perm].contiguous(), T[perm].contiguous(), batch, lr_e, WC, TAU))
        if log:
            log(f"epoch {e}: mean loss {losses[-1]:.5f} (lr {lr_e:.2e})")
    return net, losses


def save(net, path, **meta):
    np.savez(path, params=net.params(), feature_scale=np.float32(net.feature_scale()), **meta)


def load(path):
    z = np.load(path, allow_pickle=False)
    return axis.MLP(params=z["params"], feature_scale=float(z["feature_scale"]))
