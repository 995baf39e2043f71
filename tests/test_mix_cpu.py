"""The request recipes of the batching and serving workloads (CPU).

Both are deterministic in their seed, keep every image percolating along its
flow direction and stay inside the recipe's ranges; the serving recipe's
edges are log-uniform in [48, 512] and its directions and methods uniform.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from ns3d_amd import mix


def test_serving_draws_follow_the_recipe():
    draws = [mix.serving_draw(q) for q in range(4000)]
    edges = np.array([e for e, _, _ in draws])
    assert edges.min() >= 48 and edges.max() <= 512
    # log-uniform: P(edge <= e) = ln(e / 48) / ln(512 / 48)
    for e in (64, 128, 256):
        want = math.log(e / 48) / math.log(512 / 48)
        assert abs(np.mean(edges <= e) - want) < 0.03, e
    dirs = np.bincount([d for _, d, _ in draws], minlength=4)[1:]
    assert all(abs(c / len(draws) - 1 / 3) < 0.03 for c in dirs), dirs  # x, y, z a third each
    meth = {m: 0 for m in mix.METHODS}
    for _, _, m in draws:
        meth[m] += 1
    assert all(abs(c / len(draws) - 0.25) < 0.03 for c in meth.values()), meth
    assert draws == [mix.serving_draw(q) for q in range(4000)]


def test_serving_mix_sample():
    a = mix.serving_mix(6, max_edge=64)
    b = mix.serving_mix(6, max_edge=64)
    for pos, ((img, d), (img2, d2)) in enumerate(zip(a, b)):
        assert d == d2 and np.array_equal(img, img2)
        e = d["shape"][0]
        assert d["shape"] == [e] * 3 and img.shape == (e, e, e) and 48 <= e <= 64
        assert (e, d["direction"], d["method"]) == mix.serving_draw(d["recipe_index"])
        assert d["skipped_before"] == d["recipe_index"] - pos
        assert mix.percolates(img, d["direction"] - 1)
        assert 0.30 <= d["porosity_target"] <= 0.90
        assert abs(d["porosity"] - (1.0 - img.mean())) < 1e-4
        assert d["fluid"] == int((img == 0).sum())
    # the skipped requests are exactly the larger ones
    kept = {d["recipe_index"] for _, d in a}
    for q in range(a[-1][1]["recipe_index"]):
        assert (q in kept) == (mix.serving_draw(q)[0] <= 64)


def test_request_mix_sample():
    reqs = mix.request_mix(5, edges=(32, 48))
    again = mix.request_mix(5, edges=(32, 48))
    for q, ((img, d), (img2, d2)) in enumerate(zip(reqs, again)):
        assert d == d2 and np.array_equal(img, img2)
        assert all(e in (32, 48) for e in img.shape) and list(img.shape) == d["shape"]
        assert d["generator"] == ("rsa", "boolean", "fibres")[q % 3]
        assert mix.percolates(img, 0)
        assert 0.35 <= d["porosity_target"] <= 0.90


def test_serving_draw_rejects_bad_range():
    with pytest.raises(ValueError):
        mix.serving_draw(0, edge_range=(64, 48))
