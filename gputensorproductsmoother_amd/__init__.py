"""MI355X-native geometric-multigrid hot path of GPUTensorProductSmoother.

Hand-written gfx950 HIP kernels (vertex-patch Schwarz smoother with
fast-diagonalisation local solves, matrix-free Laplace, level transfer, CG
vector kernels) behind the C ABI in include/gtps.h.  The host-side mirror of
the reference's PatchSmoother / LaplaceOperator / MGTransferCUDA /
MultigridSolvers API lives in ``gputensorproductsmoother_amd.core`` (it
imports torch, so it is not imported here).
"""
from ._lib import GtpsError, load, LIB_PATH  # noqa: F401

__all__ = ["GtpsError", "load", "LIB_PATH"]
