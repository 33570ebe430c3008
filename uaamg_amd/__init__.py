"""uaamg_amd -- MI355X (gfx950) UAAMG-preconditioned CG pressure solve.

The product is the HIP library uaamg_amd/lib/libuaamg_hip.so behind the C-ABI
in include/uaamg.h; this package only loads it (ctypes) and exposes a thin
Python view used by tests and the benchmark.  There is no CPU fallback: if the
library is missing or no gfx950 device is present, calls raise.
"""
from .capi import (Lib, Context, Batch, Server, Stats, Opts, GridDesc, load, lib_path,
                   FP64, FP32, MIXED_BF16, MIXED_FP16, MIXED_BF16R, MIXED_F64, TOL_ABS, TOL_REL,
                   NEUMANN, DIRICHLET, status_name, OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_NODEVICE,
                   ERR_ADMISSION, NOT_CONVERGED, BREAKDOWN, NONFINITE,
                   FORM_ROWS, FORM_ROWS_COARSE, FORM_TAIL, FORM_PMV, FORM_XMASK, FORM_NT_X, FORM_NT_P,
                   FORM_NT_D, FORM_NT_U, FORM_SPLIT, FORM_NT, FORM_DEFAULT, PRECOND_UAAMG, PRECOND_NONE,
                   PRECOND_GMG, PRECOND_AMG, HIERARCHY_REFERENCE, HIERARCHY_KEEP_FLUID)
from .build import build, build_host_tools

__all__ = ["Lib", "Context", "Batch", "Server", "Stats", "Opts", "GridDesc", "load", "lib_path", "build", "build_host_tools",
           "FP64", "FP32", "MIXED_BF16", "MIXED_FP16", "MIXED_BF16R", "MIXED_F64", "TOL_ABS", "TOL_REL",
           "NEUMANN", "DIRICHLET", "status_name", "OK", "ERR_ARG", "ERR_HIP", "ERR_OOM", "ERR_NODEVICE",
           "ERR_ADMISSION", "NOT_CONVERGED", "BREAKDOWN", "NONFINITE",
           "FORM_ROWS", "FORM_ROWS_COARSE", "FORM_TAIL", "FORM_PMV", "FORM_XMASK", "FORM_NT_X", "FORM_NT_P",
           "FORM_NT_D", "FORM_NT_U", "FORM_SPLIT", "FORM_NT", "FORM_DEFAULT", "PRECOND_UAAMG", "PRECOND_NONE",
           "PRECOND_GMG", "PRECOND_AMG", "HIERARCHY_REFERENCE", "HIERARCHY_KEEP_FLUID"]
