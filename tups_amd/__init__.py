"""tups_amd — MI355X-native vertex hot path of the TUPS truncated-unity parquet solver.

The product is the HIP library ``libtups_gpu.so`` (sources in ``csrc/``, C ABI
in ``include/tups_gpu.h``).  This package holds its build script, the ctypes
mirror of the reference routines (:mod:`tups_amd.runtime`) and the Fortran
ISO_C_BINDING modules (``fortran/``) for the reference's own driver.
"""
from .runtime import (BSE_CLOSED_IR, BSE_CLOSED_LU, BSE_LADDER_FP64, BSE_LADDER_SPLIT, CH_D, CH_M, CH_S, CH_T, CHI,
                      F, GAMMA, LAMBDA, M, M_NAMES, PHIR, TupsBatch, TupsContext, TupsError, TupsUneven, build_tables,
                      load_library, read_prms)

__all__ = ["TupsContext", "TupsBatch", "TupsUneven", "TupsError", "load_library", "build_tables", "read_prms", "GAMMA", "F",
           "CHI", "PHIR", "M", "LAMBDA", "CH_D", "CH_M", "CH_S", "CH_T", "BSE_LADDER_FP64", "BSE_LADDER_SPLIT",
           "BSE_CLOSED_LU", "BSE_CLOSED_IR", "M_NAMES"]
