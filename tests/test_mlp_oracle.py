"""CPU checks of the proposer-network oracle (oracle/mlp_oracle.py): its
analytic gradient against central finite differences of the same float64
network, and its BF16 rounding against hand-worked values."""
import numpy as np

import mlp_oracle as M


def _net(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.N_PARAMS)
    p[M.O_W1:M.O_B1] = rng.uniform(-1, 1, M.O_B1 - M.O_W1)
    p[M.O_B1:M.O_W2] = rng.uniform(-0.1, 0.1, M.O_W2 - M.O_B1)
    p[M.O_W2:M.O_B2] = rng.uniform(-0.2, 0.2, M.O_B2 - M.O_W2)
    p[M.O_B2:M.O_W3] = rng.uniform(-0.1, 0.1, M.O_W3 - M.O_B2)
    p[M.O_W3:M.O_B3] = rng.uniform(-0.2, 0.2, M.O_B3 - M.O_W3)
    return p


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e-39], dtype=np.float32)
    got = M.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6); 1 + 2^-9 rounds down
    assert got.tolist()[:5] == [1.0, 1.0, 1.0 + 2.0 ** -6, 1.0, -2.5]
    assert got[5] > 0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _net(1)
    x = rng.normal(size=(20, 6))
    t = rng.normal(size=(20, 4))
    wc, tau = 0.7, 0.8
    _, g = M.loss_grad(p, x, t, wc, tau, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(M.O_W1, M.O_B1), 20), rng.choice(np.arange(M.O_B1, M.O_W2), 10),
                          rng.choice(np.arange(M.O_W2, M.O_B2), 40), rng.choice(np.arange(M.O_B2, M.O_W3), 10),
                          rng.choice(np.arange(M.O_W3, M.O_B3), 20), np.arange(M.O_B3, M.N_PARAMS)])
    h = 1e-6
    for j in idx:
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (M.loss_grad(pp, x, t, wc, tau, emulate=False)[0] - M.loss_grad(pm, x, t, wc, tau, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 + 1e-5 * abs(fd), (j, fd, g[j])


def test_emulated_mode_tracks_exact_mode():
    """BF16 emulation perturbs the gradient by a few 1e-3, not more."""
    rng = np.random.default_rng(2)
    p = _net(2)
    x = rng.normal(size=(64, 6))
    t = rng.normal(size=(64, 4))
    _, g64 = M.loss_grad(p, x, t, 1.0, 0.9, emulate=False)
    _, g16 = M.loss_grad(p, x, t, 1.0, 0.9, emulate=True)
    assert np.linalg.norm(g16 - g64) <= 0.05 * np.linalg.norm(g64)
