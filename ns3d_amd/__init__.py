"""MI355X implementation of NS3D's SIMPLE-family steady Navier-Stokes outer
iteration (SIMPLEC / SIMPLER / SIMPLERR / IDEAL) for Darcy permeability.

The compute path is libns3d_gpu.so (HIP kernels for gfx950 behind the C ABI in
include/ns3d.h); ``ns3d_amd.gpu`` binds it, ``ns3d_amd.host`` drives the outer
loops, ``ns3d_amd.vti`` and ``ns3d_amd.generators`` handle images.
"""
__all__ = ["gpu", "vti", "generators"]
