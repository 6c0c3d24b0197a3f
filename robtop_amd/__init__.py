"""robtop_amd — MI355X-native (gfx950) multigrid FEM elasticity solve and SIMP
loop of lavenklau/robtop, behind the reference's Grid / optimization
interface. The compute runs in hand-written HIP kernels in librobtop_amd.so
(built by ``python -m robtop_amd.build``); this package is the host-side
mirror of the reference's entry points.
"""
from ._lib import (RobtopError, RT_SOLVE_FP32, RT_SOLVE_FP64, RT_SOLVE_MIXED, RT_SOLVE_FP32_IR, RT_SOLVE_RELIABLE,
                   LIB_PATH, lib)
from .grid import Grid, HierarchyGrid, voxelize_mesh
from .optimization import (MMA, Parameter, converge_criteria, computeSensitivity, updateDensities, solveFEM,
                           ordinary_topopt, worst_case_topopt)

__all__ = ["RobtopError", "RT_SOLVE_FP32", "RT_SOLVE_FP64", "RT_SOLVE_MIXED", "RT_SOLVE_FP32_IR", "RT_SOLVE_RELIABLE",
           "LIB_PATH", "lib", "Grid",
           "HierarchyGrid", "voxelize_mesh", "Parameter", "converge_criteria", "computeSensitivity",
           "updateDensities", "solveFEM", "ordinary_topopt", "worst_case_topopt", "MMA"]
