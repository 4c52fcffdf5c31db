"""MI355X-native incompressible Navier-Stokes step (2D / 3D staggered grids).

The step runs as hand-written HIP kernels for gfx950 in ``libaerolume_amd.so``, called
through a C ABI (``include/aerolume_amd.h``). The Python layer keeps the reference's
solver API (``FlowSolver``, ``SolverConfig``, ``step_full`` ...).
"""
__version__ = "0.1.0"
