"""mgk_correlated_field_noalloc (the correlated field with every buffer beyond
the caller's in a static device array) against mgk_correlated_field, the
entry point it varies: two calls in a row, the second on a smaller grid, with
caller noise and with device normals, on a stream of the caller's.  Bar:
1e-12 of the field's scale (the filter taps are computed on the device in
FP64 instead of on the host in long double).

The test's own device buffers come from the HIP runtime the library itself
runs on (the process holds one HIP runtime: no second one is loaded)."""
import ctypes as C

import numpy as np
import pytest

from mgkmc_cpp_amd import abi, fields
from mgkmc_cpp_amd._lib import gpu
from mgkmc_cpp_amd.solver import correlated_field_device

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    gpu()   # the library loads its HIP runtime first; the same one is returned below
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def _dev(h, nbytes):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), nbytes) == 0
    return p


def _noalloc(h, stream, shape, mean, std, corr, seed, noise, clip_min=None):
    D = len(shape)
    N = int(np.prod(shape))
    out_d = _dev(h, N * 8)
    nz_d = None
    try:
        poison = np.full(N, np.nan)
        assert h.hipMemcpy(out_d, poison.ctypes.data, N * 8, H2D) == 0
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64).reshape(-1)
            nz_d = _dev(h, N * 8)
            assert h.hipMemcpy(nz_d, nz.ctypes.data, N * 8, H2D) == 0
        rc = gpu().mgk_correlated_field_noalloc(D, shape[0], shape[1], shape[2] if D == 3 else 1, float(mean),
                                                float(std), float(corr), int(clip_min is not None),
                                                float(clip_min or 0.0), 0, 0.0, int(seed), nz_d, out_d, stream)
        assert rc == abi.MGK_OK, gpu().mgk_last_error()
        assert h.hipStreamSynchronize(stream) == 0
        got = np.empty(N)
        assert h.hipMemcpy(got.ctypes.data, out_d, N * 8, D2H) == 0
        return got
    finally:
        h.hipFree(out_d)
        if nz_d is not None:
            h.hipFree(nz_d)


def test_noalloc_field_equals_the_allocating_entry_point():
    h = _hip()
    stream = C.c_void_p()
    assert h.hipStreamCreate(C.byref(stream)) == 0
    try:
        # call 1: a 3D grid with the caller's noise; call 2: a smaller 2D grid
        # with device normals (the same Philox stream keyed by the seed)
        cases = [((48, 40, 36), 4.0, 11, True, 60.0), ((24, 20), 3.0, 5, False, None)]
        for shape, corr, seed, with_noise, cmin in cases:
            N = int(np.prod(shape))
            noise = fields.box_muller_normals(seed, N) if with_noise else None
            want = correlated_field_device(shape, 85.0, 12.75, corr, seed, clip_min=cmin, noise=noise)
            got = _noalloc(h, stream, shape, 85.0, 12.75, corr, seed, noise, clip_min=cmin)
            assert np.isfinite(got).all()
            assert np.abs(got - want).max() < 1e-12 * 85.0, (shape, np.abs(got - want).max())
    finally:
        h.hipStreamDestroy(stream)
