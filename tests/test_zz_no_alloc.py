"""GPU: uaamg_amgpcg_noalloc -- the amgpcg solve with every intermediate
buffer in the library's fixed device array instead of allocations -- gives the
same x and iteration count as uaamg_amgpcg (same kernels, same order), on two
calls in a row, the second with a smaller N (the array keeps the first call's
contents, which the second must not see), in FP64 and FP32, on the default
stream and on torch's stream."""
import ctypes as C

import numpy as np
import pytest

import uaamg_amd as U

pytestmark = pytest.mark.gpu


def _rhs(N, seed):
    """sin*sin load of test_amgpcg.cpp plus seeded noise, row-major (N+2)^2, zero ghosts."""
    h = 1.0 / (N + 1)
    i = np.arange(N + 2)
    b = 2 * np.pi ** 2 * np.outer(np.sin(np.pi * i * h), np.sin(np.pi * i * h))
    b += 0.1 * np.random.default_rng(seed).standard_normal(b.shape)
    b[0, :] = b[-1, :] = 0
    b[:, 0] = b[:, -1] = 0
    return b.reshape(-1)


def _noalloc(lib, N, b, max_iter, tol, precision, stream):
    import torch
    bd = torch.from_numpy(b).to("cuda:0")
    xd = torch.full_like(bd, float("nan"))          # x is written whole, ghosts included
    torch.cuda.synchronize()
    it = C.c_int(-1)
    rc = lib.L.uaamg_amgpcg_noalloc(N, C.c_void_p(bd.data_ptr()), C.c_void_p(xd.data_ptr()), max_iter, tol,
                                    C.byref(it), precision, stream)
    torch.cuda.synchronize()
    assert rc == U.OK, U.status_name(rc)
    return xd.cpu().numpy(), it.value


@pytest.mark.parametrize("precision", [U.FP64, U.FP32])
def test_noalloc_amgpcg_matches_amgpcg(gpu, precision):
    import torch
    for k, N in enumerate((300, 97)):        # the second call smaller, odd N (odd coarse levels)
        b = _rhs(N, seed=N)
        tol = 1e-6 * float(np.linalg.norm(b))
        print(f"[noalloc] precision={precision} N={N}", flush=True)
        x_ref, it_ref = gpu.amgpcg(N, b, 200, tol, precision)
        stream = None if k == 0 else C.c_void_p(torch.cuda.current_stream().cuda_stream)
        x, it = _noalloc(gpu, N, b, 200, tol, precision, stream)
        assert it == it_ref < 200, (it, it_ref)
        np.testing.assert_array_equal(x, x_ref)


def test_noalloc_amgpcg_rejects_oversize(gpu):
    """Sizes whose buffers exceed the fixed array are refused before any device work."""
    N = 738
    b = np.zeros((N + 2) ** 2)
    x = np.zeros_like(b)
    it = C.c_int(-1)
    rc = gpu.L.uaamg_amgpcg_noalloc(N, b.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), 10, 1e-6,
                                    C.byref(it), U.FP64, None)
    assert rc == U.ERR_ADMISSION and it.value == -1
