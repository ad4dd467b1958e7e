"""MI355X (gfx950) implementation of the IMEX_SfloW2D per-timestep hot path.

The compute path is the HIP library ``libimexsw.so`` (csrc/), reached through
the C-ABI declared in include/imexsw.h; ``Solver`` is the host-side mirror of
the reference's ``solver_2d`` module interface (src/solver_2d.f90).
"""
from .params import ImexswParams, make_params  # noqa: F401
