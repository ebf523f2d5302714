"""pyquake3d_amd -- MI355X-native implementation of PyQuake3D's per-step hot path.

Layers (host Python over a C-ABI library of hand-written gfx950 HIP kernels):

* ``greenfunc``  triangular-dislocation stress and traction Green's functions
                 (drop-in for SH_greenfunction / TDstressFS_C.so)
* ``readmsh``    mesh reader and element frames (drop-in for readmsh.py)
* ``QDsim_gpu``  dense quasi-dynamic simulator with device-resident stepping
                 (drop-in for QDsim_gpu.QDsim's GPU path)
* ``synth``      synthetic workloads in the reference's input formats

The compiled library lives at ``pyquake3d_amd/libpq3d.so``; nothing here falls
back to a CPU implementation.
"""
__version__ = "0.1.0"
