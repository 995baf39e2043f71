"""The stand-alone solve without device allocation (GPU).

ns3d_bicgstab_stencil_noalloc runs the solve of ns3d_bicgstab_stencil on
caller-owned device buffers, with its workspace in one fixed per-GPU array.
Against the allocating entry point on the same inputs: bit-identical x and
the same (iterations, info, error), on two calls in a row -- a 64^3 pressure
system and then a smaller one, so the second call finds the first one's
contents in the shared workspace -- and on a non-default stream.  A zero
right-hand side gives x = 0 as in the allocating call.  Sizes beyond the
workspace, host pointers and out-of-range neighbour rows are errors.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from ns3d_amd import generators
from ns3d_amd import gpu as G

pytestmark = pytest.mark.gpu


def pressure_system(img):
    """A SIMPLEC pressure-correction system of img (matrix and right-hand
    side as the context assembles them)."""
    with G.NavierStokesGPU(img, dx=1.0, mu=1.0, rho=1e-3, p0=1.0, gradp=1.0 / img.shape[0]) as g:
        g.zero_fields()
        g.init_pressure_linear()
        g.assemble_momentum(0.8)
        g.compute_velocity_field(0.8, 0, 0.0, 1e-10, 3000)
        g.assemble_pressure(G.PC_SIMPLEC, 0, 1.0, G.SET_VEL)
        diag, off, nbr = g.matrix(G.SYS_P)[:3]
        b = g.get("b")
    return diag, off, nbr, b


def on_gpu(diag, off, nbr, b, x0):
    dev = torch.device("cuda:0")
    return (torch.from_numpy(np.ascontiguousarray(diag)).to(dev),
            torch.from_numpy(np.ascontiguousarray(off).reshape(-1)).to(dev),
            torch.from_numpy(np.ascontiguousarray(nbr, np.int32).reshape(-1)).to(dev),
            torch.from_numpy(np.ascontiguousarray(b)).to(dev),
            torch.from_numpy(np.array(x0, np.float64)).to(dev))


def both(diag, off, nbr, b, x0, tol, maxit, stream=None):
    x_ref, info_ref = G.bicgstab_stencil(diag, off, nbr, b, x0, tol, maxit)
    d, o, nb, bb, x = on_gpu(diag, off, nbr, b, x0)
    info = G.bicgstab_stencil_noalloc(d, o, nb, bb, x, tol, maxit, stream=stream)
    torch.cuda.synchronize()
    return x_ref, info_ref, x.cpu().numpy(), info


def test_noalloc_matches_allocating_entry_point_twice():
    big = generators.rsa_spheres((64, 64, 64), 8.0, 0.40, 7)[0]
    small = generators.rsa_spheres((24, 22, 20), 3.5, 0.35, 33)[0]
    for step, img in enumerate((big, small)):
        diag, off, nbr, b = pressure_system(img)
        n = diag.size
        assert n <= G.noalloc_max_rows()
        x0 = 1e-3 * np.random.default_rng(step).standard_normal(n)
        print(f"[no-alloc] call {step + 1}: {n} rows", flush=True)
        x_ref, info_ref, x, info = both(diag, off, nbr, b, x0, 1e-10, 5000)
        assert info_ref[1] == 0 and info == info_ref, (info, info_ref)
        assert np.array_equal(x, x_ref)


def test_noalloc_on_a_stream_and_zero_rhs():
    img = generators.rsa_spheres((20, 18, 16), 3.0, 0.35, 5)[0]
    diag, off, nbr, b = pressure_system(img)
    n = diag.size
    s = torch.cuda.Stream()
    print(f"[no-alloc] stream call: {n} rows", flush=True)
    x_ref, info_ref, x, info = both(diag, off, nbr, b, np.zeros(n), 1e-9, 3000, stream=s)
    assert info == info_ref and np.array_equal(x, x_ref)
    # zero right-hand side: x = 0, iterations = maxit (Eigen's contract)
    x0 = np.random.default_rng(1).standard_normal(n)
    x_ref, info_ref, x, info = both(diag, off, nbr, np.zeros(n), x0, 1e-5, 33)
    assert info == info_ref == (33, 0, 1e-5) and np.all(x == 0)


def test_noalloc_rejects_what_it_cannot_run():
    img = generators.rsa_spheres((16, 14, 12), 3.0, 0.35, 9)[0]
    diag, off, nbr, b = pressure_system(img)
    n = diag.size
    d, o, nb, bb, x = on_gpu(diag, off, nbr, b, np.zeros(n))
    lib = G.lib()
    info = G.SolveInfo()
    cap = G.noalloc_max_rows()
    # beyond the workspace: refused before anything is read
    assert lib.ns3d_bicgstab_stencil_noalloc(cap + 1, d.data_ptr(), o.data_ptr(), nb.data_ptr(), bb.data_ptr(),
                                             x.data_ptr(), 1e-8, 100, info, None) == 1
    assert "exceeds the workspace maximum" in lib.ns3d_last_error().decode()
    # host memory where device memory is expected
    hb = np.ascontiguousarray(b)
    assert lib.ns3d_bicgstab_stencil_noalloc(n, d.data_ptr(), o.data_ptr(), nb.data_ptr(), hb.ctypes.data,
                                             x.data_ptr(), 1e-8, 100, info, None) == 1
    assert "not device memory" in lib.ns3d_last_error().decode()
    # a neighbour row outside [0, n): found on the device, before the solve
    bad = nbr.copy()
    bad[2, n // 2] = n
    nb_bad = torch.from_numpy(np.ascontiguousarray(bad, np.int32).reshape(-1)).to("cuda:0")
    assert lib.ns3d_bicgstab_stencil_noalloc(n, d.data_ptr(), o.data_ptr(), nb_bad.data_ptr(), bb.data_ptr(),
                                             x.data_ptr(), 1e-8, 100, info, None) == 1
    assert "neighbour index out of range" in lib.ns3d_last_error().decode()
    # and the entry point still solves afterwards
    x_ref, info_ref, x2, info2 = both(diag, off, nbr, b, np.zeros(n), 1e-9, 3000)
    assert info2 == info_ref and np.array_equal(x2, x_ref)
