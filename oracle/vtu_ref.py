"""The reference's VTU writer restated in Python, line for line in output:
vtu_writer.hpp:22-166 (reconstruct_displacement_element_centered), :169-357
(export_to_vtu) and :359-393 (export_to_vtu_2d).

TEST INFRASTRUCTURE ONLY: the checker of the library's VTU output, for small
grids (pure Python loops).  Numbers are printed as C++ streams print a double
by default (printf "%g", six significant digits).

Pinned by: a uniform normal strain, whose reconstructed displacement is the
linear field e_aa * x_a less its x = 0 face mean (tests/test_vtu_and_logs.py).
"""
import numpy as np


def _g(x):
    return "%g" % x


def reconstruct_displacement(nx, ny, nz, eps, pixel):
    """eps: (N, 3, 3) in voxel order x*ny*nz + y*nz + z; returns (Nx*Ny*Nz, 3)."""
    Nx, Ny, Nz = nx + 1, ny + 1, nz + 1
    u = [[0.0, 0.0, 0.0] for _ in range(Nx * Ny * Nz)]

    def node(i, j, k):
        return i * Ny * Nz + j * Nz + k

    def elem(i, j, k):
        return i * ny * nz + j * nz + k

    # normal strains (:45-85)
    for i in range(1, Nx):
        for j in range(Ny):
            for k in range(Nz):
                u[node(i, j, k)][0] = u[node(i - 1, j, k)][0] + eps[elem(i - 1, min(j, ny - 1), min(k, nz - 1))][0][0] * pixel
    for j in range(1, Ny):
        for i in range(Nx):
            for k in range(Nz):
                u[node(i, j, k)][1] = u[node(i, j - 1, k)][1] + eps[elem(min(i, nx - 1), j - 1, min(k, nz - 1))][1][1] * pixel
    for k in range(1, Nz):
        for i in range(Nx):
            for j in range(Ny):
                u[node(i, j, k)][2] = u[node(i, j, k - 1)][2] + eps[elem(min(i, nx - 1), min(j, ny - 1), k - 1)][2][2] * pixel
    # shear, half the engineering shear of the cell below (:87-150)
    for j in range(1, Ny):
        for i in range(Nx):
            for k in range(Nz):
                u[node(i, j, k)][0] += 0.5 * eps[elem(min(i, nx - 1), j - 1, min(k, nz - 1))][0][1] * pixel
    for k in range(1, Nz):
        for i in range(Nx):
            for j in range(Ny):
                u[node(i, j, k)][0] += 0.5 * eps[elem(min(i, nx - 1), min(j, ny - 1), k - 1)][0][2] * pixel
    for i in range(1, Nx):
        for j in range(Ny):
            for k in range(Nz):
                u[node(i, j, k)][1] += 0.5 * eps[elem(i - 1, min(j, ny - 1), min(k, nz - 1))][0][1] * pixel
    for k in range(1, Nz):
        for i in range(Nx):
            for j in range(Ny):
                u[node(i, j, k)][1] += 0.5 * eps[elem(min(i, nx - 1), min(j, ny - 1), k - 1)][1][2] * pixel
    for i in range(1, Nx):
        for j in range(Ny):
            for k in range(Nz):
                u[node(i, j, k)][2] += 0.5 * eps[elem(i - 1, min(j, ny - 1), min(k, nz - 1))][0][2] * pixel
    for j in range(1, Ny):
        for i in range(Nx):
            for k in range(Nz):
                u[node(i, j, k)][2] += 0.5 * eps[elem(min(i, nx - 1), j - 1, min(k, nz - 1))][1][2] * pixel
    # anchor: mean of the x = 0 face (:152-163)
    u0 = [0.0, 0.0, 0.0]
    for j in range(Ny):
        for k in range(Nz):
            for a in range(3):
                u0[a] += u[node(0, j, k)][a]
    u0 = [v / (Ny * Nz) for v in u0]
    return [[p[a] - u0[a] for a in range(3)] for p in u]


def _vm(t, factor):
    tr = t[0][0] + t[1][1] + t[2][2]
    s = 0.0
    for j in range(3):          # column-major squared norm
        for i in range(3):
            d = t[i][j] - (tr / 3.0 if i == j else 0.0)
            s += d * d
    return (factor * s) ** 0.5


def export_to_vtu(nx, ny, nz, eps, sig, E, nu, pixel, eps_plastic=None, soft=None):
    """The file text of export_to_vtu; tensors (N, 3, 3), soft (N, 2)."""
    eps = [[[float(eps[n][i][j]) for j in range(3)] for i in range(3)] for n in range(len(eps))]
    sig = [[[float(sig[n][i][j]) for j in range(3)] for i in range(3)] for n in range(len(sig))]
    npts = (nx + 1) * (ny + 1) * (nz + 1)
    ncell = nx * ny * nz
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
           '  <UnstructuredGrid>\n',
           '    <Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (npts, ncell)]
    u = reconstruct_displacement(nx, ny, nz, eps, pixel)
    out.append('      <PointData Vectors="displacement">\n')
    out.append('        <DataArray type="Float64" Name="displacement" NumberOfComponents="3" format="ascii">\n')
    for p in u:
        out.append("%s %s %s\n" % (_g(p[0]), _g(p[1]), _g(p[2])))
    out.append('        </DataArray>\n')
    out.append('      </PointData>\n')
    out.append('      <CellData>\n')

    def cells(name, vals):
        out.append('        <DataArray type="Float64" Name="%s" format="ascii">\n' % name)
        for i in range(ncell):
            out.append(_g(vals[i]) + ("\n" if (i % 10 == 9 or i == ncell - 1) else " "))
        out.append('        </DataArray>\n')

    def tensor(prefix, t, factor):
        for tag, (i, j) in zip(("xx", "yy", "zz", "xy", "xz", "yz"), ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
            cells(prefix + tag, [t[n][i][j] for n in range(ncell)])
        cells(prefix + "vm", [_vm(t[n], factor) for n in range(ncell)])

    cells("E", [float(v) for v in E])
    cells("nu", [float(v) for v in nu])
    tensor("eps_", eps, 2.0 / 3.0)
    tensor("sig_", sig, 1.5)
    if eps_plastic is not None:
        ep = [[[float(eps_plastic[n][i][j]) for j in range(3)] for i in range(3)] for n in range(ncell)]
        tensor("eps_plastic_", ep, 2.0 / 3.0)
    if soft is not None:
        cells("g_p", [float(soft[n][0]) for n in range(ncell)])
        cells("g_t", [float(soft[n][1]) for n in range(ncell)])
    out.append('      </CellData>\n')
    out.append('      <Points>\n')
    out.append('        <DataArray type="Float64" Name="Points" NumberOfComponents="3" format="ascii">\n')
    for i in range(nx + 1):
        for j in range(ny + 1):
            for k in range(nz + 1):
                out.append("%s %s %s\n" % (_g(i * pixel), _g(j * pixel), _g(k * pixel)))
    out.append('        </DataArray>\n')
    out.append('      </Points>\n')
    out.append('      <Cells>\n')
    out.append('        <DataArray type="Int64" Name="connectivity" format="ascii">\n')
    NYZ, NZ = (ny + 1) * (nz + 1), nz + 1
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                n0 = i * NYZ + j * NZ + k
                n1, n2 = n0 + 1, n0 + NZ
                n3, n4 = n2 + 1, n0 + NYZ
                n5, n6 = n4 + 1, n4 + NZ
                n7 = n6 + 1
                out.append("%d %d %d %d %d %d %d %d\n" % (n0, n1, n3, n2, n4, n5, n7, n6))
    out.append('        </DataArray>\n')
    out.append('        <DataArray type="Int64" Name="offsets" format="ascii">\n')
    for i in range(1, ncell + 1):
        out.append("%d" % (i * 8) + ("\n" if (i % 10 == 0 or i == ncell) else " "))
    out.append('        </DataArray>\n')
    out.append('        <DataArray type="UInt8" Name="types" format="ascii">\n')
    for i in range(1, ncell + 1):
        out.append("12" + ("\n" if (i % 10 == 0 or i == ncell) else " "))
    out.append('        </DataArray>\n')
    out.append('      </Cells>\n')
    out.append('    </Piece>\n')
    out.append('  </UnstructuredGrid>\n')
    out.append('</VTKFile>\n')
    return "".join(out)


def export_to_vtu_2d(nx, ny, eps, sig, E, nu, pixel, eps_plastic=None, soft=None):
    """2D tensors (N, 2, 2) as the top-left block of 3D ones, nz = 1."""
    def lift(t):
        o = np.zeros((len(t), 3, 3))
        o[:, :2, :2] = np.asarray(t)
        return o
    return export_to_vtu(nx, ny, 1, lift(eps), lift(sig), E, nu, pixel,
                         lift(eps_plastic) if eps_plastic is not None else None, soft)
