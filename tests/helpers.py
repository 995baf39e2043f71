"""Shared test inputs and comparison helpers."""
from __future__ import annotations

import numpy as np

from ns3d_amd import generators


def small_image(shape=(20, 17, 13), seed=1, porosity=0.6, rmin=1.5, rmax=3.5):
    img, _, _ = generators.boolean_spheres(shape, rmin, rmax, porosity, seed)
    return img


def connected_image(shape, seed, porosity=0.62):
    """A Boolean image whose fluid is one cluster joining inlet and outlet
    (other clusters filled with solid), so every pressure system is
    non-singular."""
    from scipy import ndimage
    img = small_image(shape, seed=seed, porosity=porosity)
    lab, n = ndimage.label(img == 0)
    keep = set(np.unique(lab[0])) & set(np.unique(lab[-1])) - {0}
    assert keep, "no percolating cluster"
    sizes = {q: int((lab == q).sum()) for q in keep}
    big = max(sizes, key=sizes.get)
    return np.where(lab == big, 0, 1).astype(np.uint8)


def edge_images():
    """Degenerate and boundary cases of the index rule."""
    rng = np.random.default_rng(7)
    cases = {
        "all_fluid_4x3x5": np.zeros((4, 3, 5), np.uint8),
        "all_solid_4x3x5": np.ones((4, 3, 5), np.uint8),
        "single_fluid_voxel": np.pad(np.zeros((1, 1, 1), np.uint8), 2, constant_values=1),
        "line_1x1x9": np.zeros((1, 1, 9), np.uint8),
        "column_9x1x1": np.zeros((9, 1, 1), np.uint8),
        "slab_nz1": (rng.random((7, 6, 1)) < 0.3).astype(np.uint8),
        "random_dense": (rng.random((9, 8, 7)) < 0.5).astype(np.uint8),
        "checker": (np.indices((6, 6, 6)).sum(axis=0) % 2).astype(np.uint8),
    }
    return cases


def random_fields(ns, seed=0, zeros=0.15):
    """Random velocity/pressure vectors with mixed signs and exact zeros, so
    every branch of the upwind rule is taken."""
    rng = np.random.default_rng(seed)
    out = {}
    for name, n in ns.items():
        x = rng.standard_normal(n) * 1e-3
        x[rng.random(n) < zeros] = 0.0
        out[name] = x
    return out


def oracle_triplets(cnt, col, val):
    rows = np.repeat(np.arange(cnt.size), cnt)
    mask = np.arange(7)[None, :] < cnt[:, None]
    return rows, col[mask], val[mask]


def gpu_triplets(diag, off, nbr):
    n = diag.size
    rows = [np.arange(n)]
    cols = [np.arange(n)]
    vals = [diag]
    for d in range(6):
        present = nbr[d] != np.arange(n)
        rows.append(np.nonzero(present)[0])
        cols.append(nbr[d][present])
        vals.append(off[d][present])
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def sort_triplets(r, c, v):
    order = np.lexsort((c, r))
    return r[order], c[order], v[order]


def assert_same_matrix(orc_mat, gpu_mat, what=""):
    cnt, col, val = orc_mat
    diag, off, nbr = gpu_mat[:3]
    r1, c1, v1 = sort_triplets(*oracle_triplets(cnt, col, val))
    r2, c2, v2 = sort_triplets(*gpu_triplets(diag, off, nbr))
    assert r1.size == r2.size, f"{what}: {r1.size} vs {r2.size} stored entries"
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2), f"{what}: sparsity pattern differs"
    bad = np.nonzero((v1 + 0.0).view(np.int64) != (v2 + 0.0).view(np.int64))[0]
    assert bad.size == 0, (f"{what}: {bad.size} coefficients differ, first at row {r1[bad[0]]} col {c1[bad[0]]}: "
                           f"{v1[bad[0]]!r} vs {v2[bad[0]]!r}")


def bit_equal(a, b):
    """Identical IEEE-754 values, -0.0 and +0.0 identified."""
    a = np.asarray(a, np.float64) + 0.0
    b = np.asarray(b, np.float64) + 0.0
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def _splitmix64(z):
    """Vectorised SplitMix64 finaliser over uint64 arrays (wrapping arithmetic)."""
    z = (z + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return z ^ (z >> np.uint64(31))


def count_sketch(x, m=1024, seed=1):
    """CountSketch of x: bucket h(i) accumulates s(i) x_i with a hashed bucket
    and sign per index.  Linear in x, and E|S d|^2 = |d|^2 with relative
    standard deviation sqrt(2/m) for |S d|^2, so |S(a) - S(b)| estimates
    |a - b| to about +-2.2% (one sigma) at m = 1024 -- enough to check a
    relative L2 bar on fields too large to store."""
    x = np.asarray(x, np.float64).ravel()
    with np.errstate(over="ignore"):
        z = _splitmix64(np.arange(x.size, dtype=np.uint64) + np.uint64(seed) * np.uint64(0xD1B54A32D192ED03))
    h = (z >> np.uint64(40)) % np.uint64(m)
    s = np.where((z & np.uint64(1)) == 1, 1.0, -1.0)
    return np.bincount(h.astype(np.int64), weights=s * x, minlength=m)
