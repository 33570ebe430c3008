"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/uaamg.h declares, its structs match the ctypes mirror, and it
refuses to run without a gfx950 device (no CPU fallback)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import uaamg_amd as U
from uaamg_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "uaamg.h")
HOST_TOOL = os.path.join(ROOT, "build", "host_solver_test")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(uaamg_[a-z_]+)\s*\(", src)))


def test_header_matches_ctypes_exports():
    assert declared_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.lib_path())
    missing = [s for s in declared_functions() if not hasattr(lib, s)]
    assert not missing, missing


def test_nm_shows_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.lib_path()], capture_output=True, text=True,
                         check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for s in declared_functions():
        assert s in names, f"{s} not exported with C linkage"


_LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "uaamg.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("uaamg_grid %zu\nuaamg_opts %zu\nuaamg_stats %zu\n", sizeof(uaamg_grid), sizeof(uaamg_opts), sizeof(uaamg_stats));
  F(uaamg_grid, dim) F(uaamg_grid, n) F(uaamg_grid, h) F(uaamg_grid, solid) F(uaamg_grid, wall)
  F(uaamg_opts, precision) F(uaamg_opts, nu_pre) F(uaamg_opts, nu_post) F(uaamg_opts, coarsest_side)
  F(uaamg_opts, coarsest_sweeps) F(uaamg_opts, tol_mode) F(uaamg_opts, max_iter) F(uaamg_opts, tol)
  F(uaamg_opts, project_mean) F(uaamg_opts, fused) F(uaamg_opts, check_every) F(uaamg_opts, use_graph)
  F(uaamg_opts, timing) F(uaamg_opts, flexible) F(uaamg_opts, forms) F(uaamg_opts, xfork_level)
  F(uaamg_opts, xblocks) F(uaamg_opts, xthreads) F(uaamg_opts, rows_kc) F(uaamg_opts, march_kc_min)
  F(uaamg_opts, march_target) F(uaamg_opts, trace) F(uaamg_opts, preconditioner) F(uaamg_opts, hierarchy_mode)
  printf("UAAMG_FORM_DEFAULT %d\n", (int)UAAMG_FORM_DEFAULT);
  F(uaamg_stats, iters) F(uaamg_stats, status) F(uaamg_stats, nlevels) F(uaamg_stats, rnorm)
  F(uaamg_stats, r0norm) F(uaamg_stats, ms_total) F(uaamg_stats, ms_setup) F(uaamg_stats, ms_kernel)
  F(uaamg_stats, n_kernel) F(uaamg_stats, alg_bytes) F(uaamg_stats, kernel_set) F(uaamg_stats, forms_used)
  return 0;
}
"""


def test_struct_layout_matches_ctypes(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(_LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    c = dict(line.rsplit(" ", 1) for line in out if line)
    mirror = {"uaamg_grid": capi.GridDesc, "uaamg_opts": capi.Opts, "uaamg_stats": capi.Stats}
    for name, T in mirror.items():
        assert int(c[name]) == C.sizeof(T), name
        for fname, _ in T._fields_:
            cname = {"GridDesc": "uaamg_grid", "Opts": "uaamg_opts", "Stats": "uaamg_stats"}[T.__name__]
            assert int(c[f"{cname}.{fname}"]) == getattr(T, fname).offset, (cname, fname)
    assert int(c["UAAMG_FORM_DEFAULT"]) == U.FORM_DEFAULT


def test_default_opts_reproduce_reference():
    o = U.load().default_opts()
    assert o.precision == U.FP64
    assert (o.nu_pre, o.nu_post) == (1, 1)                     # uaamg_preconditioner.cpp:186-190
    assert o.coarsest_side == 4 and o.coarsest_sweeps == 10    # uaamg_preconditioner.cpp:20, 147-150
    assert o.tol_mode == U.TOL_ABS                             # pcg.cpp:112
    assert o.project_mean == -1 and o.flexible == 0
    assert o.forms == U.FORM_DEFAULT and o.trace == 0


def test_no_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = U.load().L
    h = C.c_void_p()
    o = U.load().default_opts()
    assert L.uaamg_create(C.byref(h), C.byref(o)) == U.ERR_NODEVICE
    with pytest.raises(RuntimeError):
        U.Context()
    with pytest.raises(RuntimeError):
        U.load().amgpcg(8, np.zeros(100), 10, 1e-8)


def test_bytes_required_scales_with_grid():
    L = U.load().L
    s = np.zeros(34 * 34 * 34, np.uint8)
    g = capi.make_grid((32, 32, 32), (1 / 32,) * 3, s)
    o64 = U.load().default_opts()
    o32 = U.load().default_opts(precision=U.FP32)
    b64 = L.uaamg_bytes_required(C.byref(g), C.byref(o64))
    b32 = L.uaamg_bytes_required(C.byref(g), C.byref(o32))
    assert 0 < b32 < b64
    assert b64 >= 34 ** 3 * 8 * 5                               # at least the five CG vectors


def test_bad_arguments_rejected():
    L = U.load().L
    assert L.uaamg_default_opts(None) == U.ERR_ARG
    assert L.uaamg_create(None, None) == U.ERR_ARG


@pytest.mark.parametrize("kw", [{"precision": U.MIXED_FP16}, {"precision": 6}, {"forms": 1 << 12},
                                {"xthreads": 32}, {"xblocks": 0}, {"march_target": 0}, {"rows_kc": -4},
                                {"tol_mode": 2}, {"nu_pre": -1}, {"preconditioner": 4},
                                {"preconditioner": -1}, {"hierarchy_mode": 2}])
def test_invalid_options_are_argument_errors(kw):
    """Options are checked before the device: an unbuilt precision (FP16
    levels) or an out-of-range form or launch shape is UAAMG_ERR_ARG, not a
    HIP error or a silent substitute."""
    L = U.load().L
    h = C.c_void_p()
    o = U.load().default_opts(**kw)
    assert L.uaamg_create(C.byref(h), C.byref(o)) == U.ERR_ARG
    b = C.c_void_p()
    o = U.load().default_opts(**{"precision": U.FP32, **kw})   # a precision the batch takes, unless under test
    assert L.uaamg_batch_create(C.byref(b), C.byref(o)) == U.ERR_ARG
    s = C.c_void_p()
    assert L.uaamg_server_create(C.byref(s), C.byref(o), 1 << 30, 4) == U.ERR_ARG


@pytest.mark.parametrize("kw", [{"forms": U.FORM_DEFAULT & ~U.FORM_TAIL}, {"preconditioner": U.PRECOND_NONE},
                                {"preconditioner": U.PRECOND_GMG}, {"preconditioner": U.PRECOND_AMG}],
                         ids=["no_tail", "identity", "gmg", "amg"])
def test_batch_and_server_need_uaamg_with_tail(kw):
    L = U.load().L
    o = U.load().default_opts(precision=U.FP32, **kw)
    b = C.c_void_p()
    assert L.uaamg_batch_create(C.byref(b), C.byref(o)) == U.ERR_ARG
    s = C.c_void_p()
    assert L.uaamg_server_create(C.byref(s), C.byref(o), 1 << 30, 4) == U.ERR_ARG


def test_bytes_required_mixed_f64_between_fp32_and_fp64():
    """FP64 CG vectors with FP32 coarse levels: more than FP32, less than FP64."""
    L = U.load().L
    s = np.zeros(66 * 66 * 66, np.uint8)
    g = capi.make_grid((64, 64, 64), (1 / 64,) * 3, s)
    b = {p: L.uaamg_bytes_required(C.byref(g), C.byref(U.load().default_opts(precision=p)))
         for p in (U.FP64, U.MIXED_F64, U.FP32)}
    assert b[U.FP32] < b[U.MIXED_F64] < b[U.FP64]


def test_bytes_required_gmg_amg_2d_only():
    """GMG / AMG exist for 2D grids only (factory_3d.cpp:21-23): the byte
    estimate of a 3D grid under them is 0 (uaamg_set_grid would refuse the
    pair), a 2D grid's equals the UAAMG context's without the tail kernel
    (the tail's stencil records belong to the UAAMG cycle alone)."""
    L = U.load().L
    g3 = capi.make_grid((32, 32, 32), (1 / 32,) * 3, np.zeros(34 ** 3, np.uint8))
    g2 = capi.make_grid((128, 96), (1 / 128, 1 / 96), np.zeros(130 * 98, np.uint8))
    b_uaamg = L.uaamg_bytes_required(C.byref(g2), C.byref(U.load().default_opts()))
    b_notail = L.uaamg_bytes_required(C.byref(g2), C.byref(U.load().default_opts(forms=U.FORM_DEFAULT & ~U.FORM_TAIL)))
    assert 0 < b_notail < b_uaamg
    for pc in (U.PRECOND_GMG, U.PRECOND_AMG):
        o = U.load().default_opts(preconditioner=pc)
        assert L.uaamg_bytes_required(C.byref(g3), C.byref(o)) == 0
        assert L.uaamg_bytes_required(C.byref(g2), C.byref(o)) == b_notail
    assert L.uaamg_bytes_required(C.byref(g3), C.byref(U.load().default_opts(preconditioner=4))) == 0


def test_host_mirror_keys():
    """include/uaamg_solver.hpp: the reference's solver keys parse and the
    factory refuses keys with no back-end here."""
    if not os.path.exists(HOST_TOOL):
        pytest.skip("host tool not built")
    r = subprocess.run([HOST_TOOL, "keys"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "keys ok" in r.stdout


def test_amgpcg_noalloc_rejects_oversize_before_the_device():
    """uaamg_amgpcg_noalloc refuses sizes whose buffers exceed its fixed
    device array (N > 737 in FP64, N > 1043 in FP32) with
    UAAMG_ERR_ADMISSION before touching any device; bad arguments first."""
    L = U.load().L
    it = C.c_int(-1)
    dummy = C.c_void_p(16)        # never dereferenced: the call is refused first
    assert L.uaamg_amgpcg_noalloc(738, dummy, dummy, 10, 1e-6, C.byref(it), U.FP64, None) == U.ERR_ADMISSION
    assert L.uaamg_amgpcg_noalloc(1044, dummy, dummy, 10, 1e-6, C.byref(it), U.FP32, None) == U.ERR_ADMISSION
    assert L.uaamg_amgpcg_noalloc(1, dummy, dummy, 10, 1e-6, C.byref(it), U.FP64, None) == U.ERR_ARG
    assert L.uaamg_amgpcg_noalloc(64, None, dummy, 10, 1e-6, C.byref(it), U.FP64, None) == U.ERR_ARG
    assert L.uaamg_amgpcg_noalloc(64, dummy, dummy, 10, 1e-6, C.byref(it), U.MIXED_BF16, None) == U.ERR_ARG
    assert it.value == -1
