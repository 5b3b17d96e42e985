"""MI355X-native exact-diagonalisation engine for the Anderson-impurity hot path.

The compute lives in libed_hip.so (hand-written HIP for gfx950, C-ABI in
include/ed_hip.h); this package is the Python host mirror of the reference's
H_class / Lanczos / Green's-function interface over that library.
"""
from .aim import AIMParams, WORKLOADS, kanamori, sector_dims, synthetic
from ._lib import BreakdownError, EDError, LIB_PATH, lib
from .hclass import (LanczosResult, Sector, Hmult__, Lanczos_fast_diagonalize, Lanczos_get_GS_sector, apply_C, apply_C_n,
                     batched_stats, compute_dynamic, compute_dynamic_batched, dynamic_from_state, delete_H, dimen_H, invert_zmtridiag, matsubara,
                     new_H, start_vector, tridiag_eig, tridiag_lowest, band_eig, SzSector, ComplexSector, ComplexSzSector)
from .block import block_lanczos

__all__ = ["AIMParams", "WORKLOADS", "kanamori", "sector_dims", "synthetic", "BreakdownError", "EDError", "LIB_PATH",
           "lib", "LanczosResult", "Sector", "Hmult__", "Lanczos_fast_diagonalize", "Lanczos_get_GS_sector",
           "apply_C", "apply_C_n", "batched_stats", "compute_dynamic", "compute_dynamic_batched", "dynamic_from_state", "delete_H", "dimen_H", "invert_zmtridiag",
           "matsubara", "new_H", "start_vector", "tridiag_eig", "tridiag_lowest", "band_eig", "block_lanczos", "SzSector",
           "ComplexSector", "ComplexSzSector"]
