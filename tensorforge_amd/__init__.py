"""gfx950-native TensorForge: batched small-GEMM and fused GEMM-chain kernels for MI355X.

The description API (``common``, ``generators``, ``frontend``) keeps the reference's module
layout; ``import tensorforge`` resolves to this package.
"""
from .interop import (get_cmake_path, get_routine_generator, get_version,  # noqa: F401
                      print_cmake_path, print_version, use_fusedgemm_cost)
