"""Sub-meshes and ownership maps for the one-rank distributed tests."""
import numpy as np

from helm_project_amd import grids


def centroid_lon(mesh):
    """Longitude of the mean vertex vector (seam-safe), degrees in [0, 360)."""
    xyz = grids.cell_centroids_xyz(mesh)
    return np.mod(np.degrees(np.arctan2(xyz[:, 1], xyz[:, 0])), 360.0)


def submesh(mesh, cells):
    """The given cells (ascending global order kept), nodes renumbered."""
    cells = np.asarray(cells, dtype=np.int64)
    off, idx = mesh["offsets"], mesh["indices"]
    lens = off[cells + 1] - off[cells]
    starts = off[cells]
    sel = np.concatenate([np.arange(s, s + n) for s, n in zip(starts, lens)]) if cells.size else np.zeros(0, np.int64)
    nodes, inv = np.unique(idx[sel], return_inverse=True)
    coords = mesh["coords"][nodes]
    new_off = np.concatenate([[0], np.cumsum(lens)])
    out = grids._mesh(coords[:, 0], coords[:, 1], new_off, inv)
    for k in ("areas", "mask"):
        if k in mesh:
            out[k] = np.asarray(mesh[k])[cells]
    return out


def band_rank(lon, n_ranks):
    return np.minimum((np.mod(lon, 360.0) / (360.0 / n_ranks)).astype(np.int64), n_ranks - 1).astype(np.int32)


def rank_share(src, dst, n_ranks, r, halo_deg):
    """Rank r's meshes: destination cells of longitude band r, source cells
    within halo_deg of the band (owned + halo), with their global ids, and
    the global owner map of the source cells."""
    s_lon, d_lon = centroid_lon(src), centroid_lon(dst)
    owner = band_rank(s_lon, n_ranks)
    width = 360.0 / n_ranks
    lo = r * width
    rel = np.mod(s_lon - (lo - halo_deg), 360.0)
    polar = np.abs(grids.cell_centroids_xyz(src)[:, 2]) > np.sin(np.radians(75.0))  # polar cells meet every longitude
    s_cells = np.nonzero((rel < width + 2 * halo_deg) | polar)[0]
    # the first local cell must be owned by r (the reference takes the rank from it)
    own = s_cells[owner[s_cells] == r]
    s_cells = np.concatenate([own[:1], np.setdiff1d(s_cells, own[:1])])
    d_cells = np.nonzero(band_rank(d_lon, n_ranks) == r)[0]
    return submesh(src, s_cells), submesh(dst, d_cells), s_cells, d_cells, owner
