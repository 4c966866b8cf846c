"""axis_batch_apply_noalloc_c (the batched apply that allocates nothing,
its long-row partial sums in a fixed per-GPU array) against
axis_batch_apply_c: the same bits, on two calls in a row, the second on a
smaller input, after a large one has left its partial sums in the array."""
import ctypes as C
import sys

import numpy as np
import oracle_py as O
import pytest

from helm_project_amd import axis, grids

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    axis.init()


def long_row_matrix(n_src, n_dst, seed):
    """COO matrix whose every 7th row has hundreds of entries (several
    segments: partial sums in the scratch array), the others 1-11."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for r in range(n_dst):
        n = min(600, n_src // 2) if r % 7 == 0 else int(rng.integers(1, 12))
        c = rng.choice(n_src, size=n, replace=False)
        rows += [r] * n
        cols += list(c)
        vals += list(rng.uniform(0.0, 1.0, n))
    return axis.matrix_from_coo(rows, cols, vals, n_src, n_dst), (rows, cols, vals)


def noalloc(m, X, Y, k, xt, yt, mode, stream=None):
    n_src, n_dst, _ = m.info()
    rc = axis.lib().axis_batch_apply_noalloc_c(m.handle, X.data_ptr(), xt, Y.data_ptr(), yt, n_src, n_dst, k, mode, stream)
    if rc != 0:
        raise axis.AxisError(axis.lib().axis_last_error_c().decode())


CASES = [(axis.PLAIN, axis.F64, axis.F64, torch.float64, torch.float64), (axis.MIXED, axis.F32, axis.F32, torch.float32, torch.float32),
         (axis.MIXED_ACC64, axis.BF16, axis.F64, torch.bfloat16, torch.float64)]


@pytest.mark.parametrize("mode,xt,yt,xdt,ydt", CASES)
def test_noalloc_equals_batch_apply(mode, xt, yt, xdt, ydt):
    say("case noalloc apply: mode", mode)
    big, big_coo = long_row_matrix(4000, 300, 1)
    small, _ = long_row_matrix(900, 60, 2)
    cfg1 = axis.generate_weights(axis.Mesh(grids.latlon_deg(2.0)), axis.Mesh(grids.cubed_sphere(24)), axis.CONSERVATIVE)
    gen = torch.Generator(device="cuda").manual_seed(7)
    for m, k in ((big, 37), (small, 9), (cfg1, 16), (big, 5)):
        n_src, n_dst, _ = m.info()
        X = torch.randn((k, n_src), generator=gen, device="cuda", dtype=torch.float64).to(xdt).contiguous()
        R = torch.empty((k, n_dst), dtype=ydt, device="cuda")
        m.batch_apply(X, R, k, xt, yt, mode)  # builds the layout; the reference result
        Y = torch.full((k, n_dst), float("nan"), dtype=ydt, device="cuda")
        noalloc(m, X, Y, k, xt, yt, mode)
        torch.cuda.synchronize()
        assert torch.equal(Y, R), (m.info(), k)
    # the FP64 plain result of the big matrix against the oracle's row loop
    if mode == axis.PLAIN:
        rows, cols, vals = big_coo
        rp, col, val = O.to_csr(rows, cols, vals, 300)
        X = torch.randn((2, 4000), generator=gen, device="cuda", dtype=torch.float64)
        Y = torch.empty((2, 300), dtype=torch.float64, device="cuda")
        big.batch_apply(X, Y, 2, xt, yt, mode)
        big.batch_apply(torch.cat([X, X]), torch.empty((4, 300), dtype=torch.float64, device="cuda"), 4, xt, yt, mode)
        noalloc(big, X, Y, 2, xt, yt, mode)
        torch.cuda.synchronize()
        for f in range(2):
            yo = O.batch_apply(rp, col, val, 4000, X[f].cpu().numpy())
            assert np.allclose(Y[f].cpu().numpy(), yo, rtol=1e-13, atol=1e-13)


def test_noalloc_refusals():
    say("case noalloc refusals")
    m, _ = long_row_matrix(3000, 100, 3)
    X = torch.zeros((9, 3000), dtype=torch.float64, device="cuda")
    Y = torch.zeros((9, 100), dtype=torch.float64, device="cuda")
    with pytest.raises(axis.AxisError, match="no apply layout"):
        noalloc(m, X, Y, 9, axis.F64, axis.F64, axis.PLAIN)
    m.batch_apply(X, Y, 9, axis.F64, axis.F64, axis.PLAIN)
    noalloc(m, X, Y, 9, axis.F64, axis.F64, axis.PLAIN)
    # a field count whose partial sums exceed the scratch array is refused
    # before anything is launched (the buffers are never touched)
    with pytest.raises(axis.AxisError, match="scratch maximum"):
        noalloc(m, X, Y, 10 ** 7, axis.F64, axis.F64, axis.PLAIN)
    torch.cuda.synchronize()
