"""Writes tests/golden/damask_32x32x1_exx1e-3.npz from the reference's DAMASK
comparison data (examples/3_uniaxial_tension_heterogenous/Damask_comparison).

Inputs, read as data only:
  E_heterogeneous_32x32x1.npy            the E field in GPa, shape (32, 32, 1),
                                         loaded with numpy.load(allow_pickle=False)
  0.001/heteroMap_load_material_inc1.vti  DAMASK's spectral solution after the
                                         one increment of 0.001/load.yaml:
                                         F_xx = 1.001, P_yy = P_zz = 0

The .vti is VTK XML ImageData with appended-inline binary arrays: each array is
base64 of a UInt32 header [n_blocks, block_size, last_block_size,
compressed_size_0 ..] followed by base64 of the zlib-compressed blocks.  Cell
data are stored x fastest (VTK order); the fixture stores them in this
package's voxel order, x slowest: voxel (x, y, z) at x*ny*nz + y*nz + z
(main.cpp:548 indexes the same way).

Output arrays:
  E        (1024,)      Pa (GPa values x 1e9, as main.cpp:137-141 scales them)
  sigma    (1024, 3, 3) DAMASK's Cauchy stress, Pa
  F        (1024, 3, 3) deformation gradient
  P        (1024, 3, 3) first Piola-Kirchhoff stress, Pa
  pixel 0.5, nu 0.3 (plane_stress_3d_32_DAMASK_comparison.yaml), F_xx 1.001

Run from the repository root:  python tests/golden/make_damask_fixture.py
"""
import base64
import os
import re
import struct
import sys
import zlib

import numpy as np

SRC = "/root/reference/examples/3_uniaxial_tension_heterogenous/Damask_comparison"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "damask_32x32x1_exx1e-3.npz")
NX, NY, NZ = 32, 32, 1


def _b64_len(nbytes):
    return 4 * ((nbytes + 2) // 3)


def decode_array(text, dtype, ncomp):
    """One binary DataArray body: header, then zlib blocks."""
    text = "".join(text.split())
    first = base64.b64decode(text[:8])
    nblocks = struct.unpack("<I", first[:4])[0]
    hlen = _b64_len(4 * (3 + nblocks))
    header = struct.unpack("<%dI" % (3 + nblocks), base64.b64decode(text[:hlen]))
    sizes = header[3:]
    data = base64.b64decode(text[hlen:])
    out = bytearray()
    off = 0
    for s in sizes:
        out += zlib.decompress(data[off:off + s])
        off += s
    expect = header[1] * (nblocks - 1) + (header[2] if header[2] else header[1])
    if len(out) != expect:
        raise ValueError("decoded %d bytes, header says %d" % (len(out), expect))
    return np.frombuffer(bytes(out), dtype=dtype).reshape(-1, ncomp)


def read_cell_arrays(path):
    xml = open(path, "r", encoding="latin1").read()
    cell = xml[xml.index("<CellData"):xml.index("</CellData>")]
    arrays = {}
    for m in re.finditer(r'<DataArray ([^>]*)>(.*?)</DataArray>', cell, re.S):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        dt = {"Float32": "<f4", "Float64": "<f8", "Int32": "<i4", "Int64": "<i8"}[attrs["type"]]
        arrays[attrs["Name"]] = decode_array(m.group(2), dt, int(attrs.get("NumberOfComponents", 1)))
    return arrays


def vtk_to_voxel_order(a):
    """(nx*ny*nz, k) with x fastest -> x slowest."""
    k = a.shape[1]
    return a.reshape(NZ, NY, NX, k).transpose(2, 1, 0, 3).reshape(NX * NY * NZ, k)


def main():
    E = np.load(os.path.join(SRC, "E_heterogeneous_32x32x1.npy"), allow_pickle=False)
    assert E.shape == (NX, NY, NZ)
    # the DAMASK geometry's per-cell Young's modulus, reordered, is the .npy
    geom = read_cell_arrays(os.path.join(SRC, "0.001", "heteroMap.vti"))
    E_geom = vtk_to_voxel_order(geom["youngsModulus"])[:, 0]
    if not np.allclose(E_geom, E.reshape(-1) * 1e9, rtol=1e-12, atol=0):
        raise ValueError("heteroMap.vti youngsModulus does not match E_heterogeneous_32x32x1.npy")
    arrays = read_cell_arrays(os.path.join(SRC, "0.001", "heteroMap_load_material_inc1.vti"))
    # a 3x3 tensor is stored row by row: sigma = P F^T / det F holds per cell
    # to the Float32 rounding of the file in that reading
    sig = vtk_to_voxel_order(arrays["phase/mechanical/sigma / Pa"].astype(np.float64)).reshape(-1, 3, 3)
    F = vtk_to_voxel_order(arrays["phase/mechanical/F / 1"].astype(np.float64)).reshape(-1, 3, 3)
    P = vtk_to_voxel_order(arrays["phase/mechanical/P / Pa"].astype(np.float64)).reshape(-1, 3, 3)
    push = np.einsum("nij,nkj->nik", P, F) / np.linalg.det(F)[:, None, None]
    if np.abs(push - sig).max() > 1e-5 * np.abs(sig).max():
        raise ValueError("sigma != P F^T / J: tensor layout not row by row")
    np.savez_compressed(OUT, E=E.reshape(-1) * 1e9, sigma=sig, F=F, P=P, pixel=0.5, nu=0.3, F_xx=1.001)
    print("wrote", OUT, "mean sigma_xx = %.6f MPa" % (sig[:, 0, 0].mean() / 1e6))


if __name__ == "__main__":
    sys.exit(main())
