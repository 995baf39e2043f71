"""Seeded synthetic porous-medium images (uint8, 1 = solid, 0 = fluid).

* ``sphere_lattice`` restates the reference's test-image generator
  (gen_spheres.cxx:60-99): a regular lattice of non-overlapping spheres;
  ``sphere_lattice(128, 12, 5)`` and ``sphere_lattice(64, 8, 3)`` rebuild
  data/spheres_128.vti and data/spheres_64.vti voxel for voxel.
* ``rsa_spheres``: random sequential addition of equal spheres with a
  counter-based PRNG (SplitMix64), for the 64^3 / 128^3 benchmark packs.
  Sequential addition of hard spheres jams near a solid fraction of 0.38; when
  the target is not reached after ``max_reject`` consecutive rejections the
  allowed overlap grows by one voxel and addition continues (the achieved
  solid fraction and the final overlap are returned).
* ``boolean_spheres``: overlapping spheres with uniform radii (Boolean model),
  added until the target porosity is reached.
* ``fibres``: random straight cylinders with isotropic axes.

Every generator is deterministic for a given seed; a voxel (i, j, k) is solid
iff its centre lies inside an inclusion, i.e. (i-cx)^2+(j-cy)^2+(k-cz)^2 <= r^2
for spheres as in gen_spheres.cxx:90-95.
"""
from __future__ import annotations

import numpy as np

_MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 counter-based generator (Steele, Lea & Flood 2014)."""

    def __init__(self, seed: int):
        self.state = seed & _MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))


def _stamp_sphere(img: np.ndarray, cx: float, cy: float, cz: float, r: float) -> int:
    """Set the voxels whose centre lies in the sphere; returns how many were fluid."""
    n0, n1, n2 = img.shape
    lo = [max(0, int(np.floor(c - r))) for c in (cx, cy, cz)]
    hi = [min(n - 1, int(np.ceil(c + r))) for c, n in zip((cx, cy, cz), (n0, n1, n2))]
    if lo[0] > hi[0] or lo[1] > hi[1] or lo[2] > hi[2]:
        return 0
    i = np.arange(lo[0], hi[0] + 1, dtype=np.float64)[:, None, None]
    j = np.arange(lo[1], hi[1] + 1, dtype=np.float64)[None, :, None]
    k = np.arange(lo[2], hi[2] + 1, dtype=np.float64)[None, None, :]
    d2 = (i - cx) * (i - cx) + (j - cy) * (j - cy) + (k - cz) * (k - cz)
    sub = img[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1]
    inside = d2 <= r * r
    added = int(np.count_nonzero(inside & (sub == 0)))
    sub[inside] = 1
    return added


def sphere_lattice(n: int, spacing: int, radius: int) -> np.ndarray:
    """Regular sphere lattice, gen_spheres.cxx:60-99 (MakeSphereLattice)."""
    img = np.zeros((n, n, n), dtype=np.uint8)
    nper = n // spacing
    r2 = float(radius) * radius
    for m in range(nper):
        for q in range(nper):
            for p in range(nper):
                cx = spacing // 2 + spacing * m
                cy = spacing // 2 + spacing * q
                cz = spacing // 2 + spacing * p
                lx, hx = max(cx - radius, 0), min(cx + radius, n - 1)
                ly, hy = max(cy - radius, 0), min(cy + radius, n - 1)
                lz, hz = max(cz - radius, 0), min(cz + radius, n - 1)
                i = np.arange(lx, hx + 1, dtype=np.float64)[:, None, None]
                j = np.arange(ly, hy + 1, dtype=np.float64)[None, :, None]
                k = np.arange(lz, hz + 1, dtype=np.float64)[None, None, :]
                d2 = (i - cx) * (i - cx) + (j - cy) * (j - cy) + (k - cz) * (k - cz)
                img[lx:hx + 1, ly:hy + 1, lz:hz + 1][d2 <= r2] = 1
    return img


def rsa_spheres(shape, radius: float, solid_fraction: float, seed: int,
                max_reject: int = 20000, max_overlap: float | None = None):
    """Random sequential addition of spheres of one radius.

    Centres are uniform in the box (spheres may be cut by the faces).  A new
    sphere is accepted when its centre is at least ``2*radius - overlap`` from
    every accepted centre; ``overlap`` starts at 0 and grows by one voxel after
    ``max_reject`` consecutive rejections.  Returns (image, achieved solid
    fraction, final overlap, number of spheres).
    """
    shape = tuple(int(s) for s in shape)
    img = np.zeros(shape, dtype=np.uint8)
    rng = SplitMix64(seed)
    centres = np.zeros((0, 3))
    overlap = 0.0
    if max_overlap is None:
        max_overlap = 2.0 * radius
    total = float(np.prod(shape))
    solid = 0
    rejects = 0
    while solid / total < solid_fraction:
        c = np.array([rng.uniform() * shape[0], rng.uniform() * shape[1], rng.uniform() * shape[2]])
        dmin = 2.0 * radius - overlap
        if centres.shape[0]:
            d2 = ((centres - c) ** 2).sum(axis=1)
            ok = bool((d2 >= dmin * dmin).all())
        else:
            ok = True
        if ok:
            centres = np.vstack([centres, c])
            solid += _stamp_sphere(img, c[0], c[1], c[2], radius)
            rejects = 0
        else:
            rejects += 1
            if rejects >= max_reject:
                if overlap >= max_overlap:
                    break
                overlap += 1.0
                rejects = 0
    return img, solid / total, overlap, int(centres.shape[0])


def boolean_spheres(shape, rmin: float, rmax: float, porosity: float, seed: int):
    """Boolean model: overlapping spheres, radii uniform in [rmin, rmax],
    added until the fluid fraction drops to ``porosity``.  Returns
    (image, achieved porosity, number of spheres)."""
    shape = tuple(int(s) for s in shape)
    img = np.zeros(shape, dtype=np.uint8)
    rng = SplitMix64(seed)
    total = float(np.prod(shape))
    count = 0
    solid = 0
    while 1.0 - solid / total > porosity:
        c = [rng.uniform() * shape[0], rng.uniform() * shape[1], rng.uniform() * shape[2]]
        r = rmin + (rmax - rmin) * rng.uniform()
        solid += _stamp_sphere(img, c[0], c[1], c[2], r)
        count += 1
    return img, 1.0 - solid / total, count


def fibres(shape, radius: float, length: float, porosity: float, seed: int):
    """Random straight cylinders (isotropic axis, uniform centre) until the
    fluid fraction drops to ``porosity``.  Returns (image, porosity, count)."""
    shape = tuple(int(s) for s in shape)
    img = np.zeros(shape, dtype=np.uint8)
    rng = SplitMix64(seed)
    total = float(np.prod(shape))
    count = 0
    solid = 0
    while 1.0 - solid / total > porosity:
        c = np.array([rng.uniform() * shape[0], rng.uniform() * shape[1], rng.uniform() * shape[2]])
        zc = 2.0 * rng.uniform() - 1.0
        ph = 2.0 * np.pi * rng.uniform()
        s = np.sqrt(max(0.0, 1.0 - zc * zc))
        d = np.array([s * np.cos(ph), s * np.sin(ph), zc])
        half = 0.5 * length
        ends = np.stack([c - half * d, c + half * d])
        lo = np.maximum(np.floor(ends.min(axis=0) - radius).astype(int), 0)
        hi = np.minimum(np.ceil(ends.max(axis=0) + radius).astype(int), np.array(shape) - 1)
        if (lo <= hi).all():
            i = np.arange(lo[0], hi[0] + 1, dtype=np.float64)[:, None, None] - c[0]
            j = np.arange(lo[1], hi[1] + 1, dtype=np.float64)[None, :, None] - c[1]
            k = np.arange(lo[2], hi[2] + 1, dtype=np.float64)[None, None, :] - c[2]
            t = i * d[0] + j * d[1] + k * d[2]
            tc = np.clip(t, -half, half)
            q2 = (i - tc * d[0]) ** 2 + (j - tc * d[1]) ** 2 + (k - tc * d[2]) ** 2
            sub = img[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1]
            inside = q2 <= radius * radius
            solid += int(np.count_nonzero(inside & (sub == 0)))
            sub[inside] = 1
        count += 1
    return img, 1.0 - solid / total, count
