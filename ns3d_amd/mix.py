"""Seeded request mixes for the batched and served forms.

A mix follows the batching workload's recipe: per image, edges drawn
independently per axis from a set (anisotropic boxes allowed), porosity
uniform in a range, and the microstructure generator split evenly among
monodisperse sphere packs (random sequential addition, radius uniform in
[4, 12]), polydisperse overlapping spheres (Boolean model, radii uniform in
[3, 15]) and random straight fibres (radius 3, length 40, isotropic axis),
with the recipe's radii whatever the box.  Image q uses seed base_seed + q.  An image whose fluid
does not connect the inlet to the outlet along the flow axis has no
permeability to compute, so it is drawn again with the next seed (the seed
used is recorded).
"""
from __future__ import annotations

import numpy as np

from . import generators
from .generators import SplitMix64


def percolates(img: np.ndarray, axis: int = 0) -> bool:
    """True if the fluid (img == 0) connects the first and last planes along
    axis (6-connectivity)."""
    from scipy import ndimage
    lab, n = ndimage.label(img == 0)
    if n == 0:
        return False
    a = np.moveaxis(lab, axis, 0)
    first = set(np.unique(a[0]).tolist()) - {0}
    last = set(np.unique(a[-1]).tolist()) - {0}
    return bool(first & last)


def make_image(kind: str, shape, porosity: float, seed: int):
    """One image of the given generator kind ('rsa', 'boolean', 'fibres');
    returns (image, achieved porosity, description)."""
    rng = SplitMix64(seed ^ 0x9E3779B97F4A7C15)
    if kind == "rsa":
        r = 4.0 + 8.0 * rng.uniform()
        img, solid, overlap, count = generators.rsa_spheres(shape, r, 1.0 - porosity, seed, max_reject=2000)
        return img, 1.0 - solid, dict(generator="rsa", radius=round(r, 3), spheres=count, overlap=overlap)
    if kind == "boolean":
        img, phi, count = generators.boolean_spheres(shape, 3.0, 15.0, porosity, seed)
        return img, phi, dict(generator="boolean", rmin=3.0, rmax=15.0, spheres=count)
    if kind == "fibres":
        img, phi, count = generators.fibres(shape, 3.0, 40.0, porosity, seed)
        return img, phi, dict(generator="fibres", radius=3.0, length=40.0, fibres=count)
    raise ValueError(kind)


def request_mix(count: int, base_seed: int = 20261203, edges=(32, 48, 64, 96), porosity=(0.35, 0.90),
                max_tries: int = 20):
    """count requests: [(image, description)], deterministic in base_seed."""
    kinds = ("rsa", "boolean", "fibres")
    out = []
    for q in range(count):
        seed = base_seed + q
        for attempt in range(max_tries):
            rng = SplitMix64(seed)
            shape = tuple(int(edges[int(rng.uniform() * len(edges)) % len(edges)]) for _ in range(3))
            phi = porosity[0] + (porosity[1] - porosity[0]) * rng.uniform()
            img, got, desc = make_image(kinds[q % 3], shape, phi, seed)
            if percolates(img, 0):
                desc.update(shape=list(shape), porosity_target=round(phi, 4), porosity=round(float(got), 4),
                            seed=seed, fluid=int((img == 0).sum()))
                out.append((img, desc))
                break
            seed += 100003  # redraw
        else:
            raise RuntimeError(f"no percolating image for request {q} after {max_tries} draws")
    return out
