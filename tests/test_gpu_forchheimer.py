"""Forchheimer sweep on the GPU: one image at four pressure gradients, run
together on one engine.  Checks what the physics fixes: the apparent
permeability falls as inertia grows, the fitted k1 is the creeping-flow
(Darcy) permeability, the inertial coefficient is finite and positive, the
quadratic law fits the low-Reynolds points, and the batched points equal the
same solves run one by one, bit for bit."""
from __future__ import annotations

import math

import pytest

from ns3d_amd import forchheimer, generators

pytestmark = pytest.mark.gpu

BASE = dict(dx=1.0, mu=1.0, rho=1.0, p0=100.0, alpha=0.8, beta=1.0, direction=1, border=0, method="SIMPLEC",
            simpletol=1e-7, maxitSIMPLE=800, vsolvertol=1e-10, psolvertol=1e-10, pcsolvertol=1e-6,
            maxitsolver=3000, device=0)


def test_forchheimer_sweep_and_fit():
    img = generators.rsa_spheres((32, 24, 24), 4.0, 0.3, 11)[0]
    g = [1e-3, 0.03, 0.1, 0.3]
    r = forchheimer.sweep(img, g, **BASE)
    pts = r["points"]
    assert all(p["converged"] == 1 for p in pts), pts
    ka = [p["K_apparent"] for p in pts]
    assert all(a > b for a, b in zip(ka, ka[1:])), ka           # inertia lowers the apparent K
    assert pts[-1]["Re"] > 0.1 and pts[0]["Re"] < 1e-2
    f = r["fit"]
    assert f["points"] == 4 and f["rel_residual"] < 1e-4, f
    assert f["k1"] == pytest.approx(pts[0]["K"], rel=2e-3), f   # Darcy limit
    assert 0 < f["k2"] < math.inf, f
    seq = forchheimer.sweep(img, [g[0], g[-1]], batch=False, **BASE)["points"]
    for a, b in zip(seq, (pts[0], pts[-1])):
        assert a["K"] == b["K"] and a["iterations"] == b["iterations"], (a, b)
