"""The C-ABI libraries load and export what the headers declare (CPU only:
no compute call is made)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ns3d_[a-z_0-9]+)\s*\(", txt)))


def test_gpu_library_exports_every_declared_symbol():
    from ns3d_amd import gpu
    path = gpu.LIB_PATH
    if not os.path.exists(path):
        pytest.fail(f"{path} not built")
    lib = ctypes.CDLL(path)
    names = declared("ns3d.h")
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(gpu.EXPORTS) == names


def test_prepare_image_runs_on_host():
    """ns3d_prepare_image is pure host code (NSBase.cxx:2425-2492)."""
    import numpy as np
    from ns3d_amd import gpu
    img = (np.arange(60).reshape(3, 4, 5) % 3 == 0).astype(np.uint8)
    r = gpu.prepare_image(img, 2, 1)
    assert r.shape == (6, 3, 5)
    assert np.array_equal(r[1:5], img[::-1].transpose(1, 0, 2))
    assert np.array_equal(gpu.prepare_image(img, 3, 0), img.transpose(2, 1, 0))
