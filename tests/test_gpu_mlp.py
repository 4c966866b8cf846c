"""The proposer network's GPU kernels (BF16 MFMA forward / backward, Adam)
against the CPU oracle oracle/mlp_oracle.py, which rounds to BF16 where the
kernels do; differences come only from FP32 summation order (and the rare
BF16 rounding or ReLU decision it flips)."""
import numpy as np
import pytest

import mlp_oracle as M
from helm_project_amd import axis

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.normal(size=(n, 3)), rng.uniform(0.2, 2.0, size=(n, 3))], axis=1).astype(np.float32)
    t = np.concatenate([0.1 * rng.normal(size=(n, 3)), np.log(x[:, 3:4] + x[:, 4:5])], axis=1).astype(np.float32)
    return x, t


@pytest.mark.parametrize("n", [1, 45, 1000])
def test_forward_matches_oracle(n):
    net = axis.MLP(seed=7, feature_scale=1.0)
    p = net.params()
    x, _ = _data(n, 3)
    out = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    net.forward(torch.tensor(x, device="cuda"), out)
    o_ref, _ = M.forward(p, x, emulate=True)
    err = np.abs(out.cpu().numpy() - o_ref).max()
    print("forward max abs err", err, "scale", np.abs(o_ref).max(), flush=True)
    assert err <= 2e-3 * max(1.0, np.abs(o_ref).max())
    net.close()


@pytest.mark.parametrize("n,wc,tau", [(32, 1.0, 0.5), (333, 0.3, 0.9)])
def test_gradient_matches_oracle(n, wc, tau):
    net = axis.MLP(seed=11, feature_scale=1.0)
    p = net.params()
    x, t = _data(n, 5)
    g = torch.empty(axis.N_MLP_PARAMS, dtype=torch.float32, device="cuda")
    loss = net.gradient(torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda"), g, wc=wc, tau=tau)
    l_ref, g_ref = M.loss_grad(p, x, t, wc, tau, emulate=True)
    g = g.cpu().numpy()
    assert loss == pytest.approx(l_ref, rel=1e-4)
    for name, a, b in (("W1", M.O_W1, M.O_B1), ("b1", M.O_B1, M.O_W2), ("W2", M.O_W2, M.O_B2), ("b2", M.O_B2, M.O_W3),
                       ("W3", M.O_W3, M.O_B3), ("b3", M.O_B3, M.N_PARAMS)):
        rel = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-30)
        print(name, "rel err", rel, flush=True)
        assert rel <= 2e-3, name
    net.close()


def test_training_reduces_loss_and_is_deterministic():
    x, t = _data(8192, 9)
    xd, td = torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda")
    runs = []
    for _ in range(2):
        net = axis.MLP(seed=3, feature_scale=1.0)
        g = torch.empty(axis.N_MLP_PARAMS, dtype=torch.float32, device="cuda")
        l0 = net.gradient(xd, td, g, wc=1.0, tau=0.5)
        losses = [net.train(xd, td, batch=1024, lr=3e-3, wc=1.0, tau=0.5) for _ in range(30)]
        l1 = net.gradient(xd, td, g, wc=1.0, tau=0.5)
        print("loss", l0, "->", l1, losses[:3], losses[-3:], flush=True)
        assert l1 < 0.25 * l0
        runs.append(net.params())
        net.close()
    assert np.array_equal(runs[0], runs[1])
