"""The C-ABI libraries load and export every function include/*.h declares
(no device call is made: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "mgkmc_cpp_amd", "lib")


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:\w+\s+)*\w+[\s\*]+(mgk_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("mgkmc_gpu.h", "libmgkmc_gpu.so"), ("mgkmc_host.h", "libmgkmc_host.so")])
def test_every_declared_symbol_is_exported(header, lib):
    path = os.path.join(LIB, lib)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (python -m mgkmc_cpp_amd.build)")
    names = _declared(header)
    assert len(names) > 10
    so = ctypes.CDLL(path)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_abi_version_and_error_path_without_device():
    from mgkmc_cpp_amd import _lib
    g = _lib.gpu()
    assert g.mgk_abi_version() == 1
    # argument validation happens before any device call
    rc = g.mgk_create(4, 8, 8, 1, 1.0, 0, 0, None, None, ctypes.byref(ctypes.c_void_p()))
    assert rc == 1 and b"null" in g.mgk_last_error()


def test_noalloc_field_rejects_grids_over_its_limit_without_device():
    from mgkmc_cpp_amd import _lib
    g = _lib.gpu()
    # 160^3 voxels exceed the 2^21-voxel scratch array; an axis over 2048 too
    rc = g.mgk_correlated_field_noalloc(3, 160, 160, 160, 85.0, 12.75, 4.0, 0, 0.0, 0, 0.0, 1, None,
                                        ctypes.c_void_p(1), None)
    assert rc == 1 and b"no-allocation limit" in g.mgk_last_error()
    rc = g.mgk_correlated_field_noalloc(2, 4096, 2, 1, 85.0, 12.75, 4.0, 0, 0.0, 0, 0.0, 1, None,
                                        ctypes.c_void_p(1), None)
    assert rc == 1 and b"no-allocation limit" in g.mgk_last_error()
    rc = g.mgk_correlated_field_noalloc(2, 8, 8, 1, 85.0, 12.75, 4.0, 0, 0.0, 0, 0.0, 1, None, None, None)
    assert rc == 1 and b"null" in g.mgk_last_error()


def test_footprint_estimate_without_device():
    # pure arithmetic: per-voxel catalogs add C*N*M doubles over a shared one,
    # the KMC state grows with M, and bad shapes are rejected
    from mgkmc_cpp_amd import _lib
    g = _lib.gpu()
    n3 = 32 * 32 * 32
    shared = g.mgk_estimate_bytes(3, 32, 32, 32, 12, 0)
    per_voxel = g.mgk_estimate_bytes(3, 32, 32, 32, 12, 1)
    assert per_voxel - shared == 6 * n3 * 12 * 8 - 6 * 12 * 8
    assert g.mgk_estimate_bytes(3, 32, 32, 32, 0, 0) < shared < g.mgk_estimate_bytes(3, 32, 32, 32, 24, 0)
    assert g.mgk_estimate_bytes(2, 64, 64, 1, 6, 0) < g.mgk_estimate_bytes(3, 16, 16, 16, 6, 0) * 16
    assert g.mgk_estimate_bytes(4, 8, 8, 8, 6, 0) == -1
