"""Sampled-row parity helpers for the full-size workloads: the oracle is run
on a destination mesh made of a sample of the destination cells (a cell's row
depends on that cell and the whole source only), and the library's rows for
those cells are compared with the oracle's."""
import numpy as np


def sample_cells(n_cells, size, seed, stride_extra=None):
    rng = np.random.default_rng(seed)
    s = rng.choice(n_cells, size=min(size, n_cells), replace=False)
    if stride_extra:
        s = np.concatenate([s, np.arange(0, n_cells, stride_extra)])
    return np.unique(s)


def sub_mesh(mesh, cells):
    """The mesh restricted to `cells` (same node array, cells in the given order)."""
    off, idx = mesh["offsets"], mesh["indices"]
    counts = off[cells + 1] - off[cells]
    return {"coords": mesh["coords"], "offsets": np.concatenate([[0], np.cumsum(counts)]).astype(np.int64),
            "indices": np.concatenate([idx[off[c]:off[c + 1]] for c in cells]).astype(np.int64)}


def check_rows(csr, ref, cells, rel=1e-12, exact=False):
    """Rows `cells` of csr = (row_ptr, col, val) against the oracle's result on
    sub_mesh(dst, cells): columns bit-exact, weights within rel (or bit-exact)."""
    rp, col, val = csr
    for k, r in enumerate(cells):
        a, b = rp[r], rp[r + 1]
        ra, rb = ref["row_ptr"][k], ref["row_ptr"][k + 1]
        assert np.array_equal(col[a:b], ref["col"][ra:rb]), f"row {r}: columns differ"
        w = ref["val"][ra:rb]
        if exact:
            assert np.array_equal(val[a:b], w), f"row {r}: weights differ"
        else:
            assert np.all(np.abs(val[a:b] - w) <= rel * np.maximum(np.abs(w), 1e-13)), f"row {r}: weights differ"


def rows_csr(ref):
    """(row_ptr, col, val) of the oracle's sampled-row result."""
    return ref["row_ptr"], ref["col"], ref["val"]
