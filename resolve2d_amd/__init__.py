"""resolve2d's step (Solver.process) on MI355X: HIP kernels behind a C ABI.

    from resolve2d_amd.solver import Solver, BodyOptions, RectangleOptions
"""
from . import _native  # noqa: F401

__all__ = ["solver", "engine", "scenes"]
