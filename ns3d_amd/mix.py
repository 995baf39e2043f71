"""Seeded request mixes for the batched and served forms.

request_mix: the batching workload.
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

serving_mix: the serving workload (see its docstring).
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


METHODS = ("SIMPLEC", "SIMPLER", "SIMPLERR", "IDEAL")


def serving_draw(q: int, base_seed: int = 20261203, edge_range=(48, 512)):
    """Edge, flow direction and method of the serving recipe's request q
    (drawn before its image, so a bounded sample can skip large requests
    without generating them)."""
    lo, hi = float(edge_range[0]), float(edge_range[1])
    if not (0 < lo <= hi):
        raise ValueError("edge_range must be 0 < lo <= hi")
    rng = SplitMix64((base_seed + q) ^ 0x5DEECE66D)
    edge = int(round(np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * rng.uniform())))
    edge = min(max(edge, int(lo)), int(hi))
    direction = 1 + int(rng.uniform() * 3) % 3
    method = METHODS[int(rng.uniform() * 4) % 4]
    return edge, direction, method


def serving_mix(count: int, base_seed: int = 20261203, edge_range=(48, 512), porosity=(0.30, 0.90),
                radii=(6.0, 20.0), max_edge: int | None = None, max_tries: int = 20):
    """The serving workload's requests: request q (seed base_seed + q) is a
    cubic image whose edge is log-uniform in edge_range, drawn from the
    precision workload's generator (Boolean model of overlapping spheres,
    radii uniform in radii) at a porosity uniform in the range, with a flow
    direction uniform in {1, 2, 3} (x, y, z) and a method uniform among
    SIMPLEC, SIMPLER, SIMPLERR and IDEAL.  An image whose fluid does not
    connect inlet and outlet along its flow direction is drawn again with
    the next seed, as in request_mix.

    max_edge bounds a sample: requests of the recipe whose edge exceeds it
    are skipped (their count is recorded as ``skipped_before``), the others
    kept in the recipe's order.  Returns count (image, description) pairs;
    the description's ``recipe_index`` is q."""
    out, skipped, q = [], 0, 0
    while len(out) < count:
        seed = base_seed + q
        edge, direction, method = serving_draw(q, base_seed, edge_range)
        if max_edge is not None and edge > max_edge:
            skipped += 1
            q += 1
            continue
        for attempt in range(max_tries):
            r2 = SplitMix64(seed)
            phi = porosity[0] + (porosity[1] - porosity[0]) * r2.uniform()
            img, got, count_s = generators.boolean_spheres((edge, edge, edge), radii[0], radii[1], phi, seed)
            if percolates(img, direction - 1):
                out.append((img, dict(recipe_index=q, seed=seed, shape=[edge] * 3, generator="boolean",
                                      rmin=radii[0], rmax=radii[1], spheres=count_s, porosity_target=round(phi, 4),
                                      porosity=round(float(got), 4), direction=direction, method=method,
                                      fluid=int((img == 0).sum()), skipped_before=skipped)))
                break
            seed += 100003  # redraw
        else:
            raise RuntimeError(f"no percolating image for serving request {q} after {max_tries} draws")
        q += 1
    return out
