"""drape_amd — MI355X-native implicit-Euler Newton–PCG cloth step behind the reference's Solver surface.

The compute path is libdrape_hip.so (hand-written HIP for gfx950 + C++ host, C-ABI in include/drape_hip.h);
this package is its Python mirror (`HipSolver`, `make_solver`) plus the benchmark scene builders (`scenes`).
"""
from .solver import (HipSolver, bake_collision_field, bsr_spmv, default_settings, eval_terms, field_query,
                     hip_available, make_solver, vertex_normals)

__all__ = ["HipSolver", "make_solver", "hip_available", "default_settings", "bsr_spmv", "eval_terms", "field_query",
           "vertex_normals", "bake_collision_field"]
