// vtu_writer.hpp -- drop-in replacement of the reference's vtu_writer.hpp:
// reconstruct_displacement_element_centered (vtu_writer.hpp:22-166),
// export_to_vtu (:169-357) and export_to_vtu_2d (:359-393) with the same
// names and arguments on the mgkt tensor types, writing the same ASCII VTK
// UnstructuredGrid (hexahedral cells, nodal displacement, cell fields; default
// stream formatting, 10 values per line for cell arrays).
#pragma once
#include <algorithm>
#include <cmath>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mgk_tensor.hpp"

// Nodal displacement of a grid of strained cells (vtu_writer.hpp:22-166).
// Component a of node n: the normal strain e_aa integrated cell by cell along
// +a from the a = 0 face, plus, for each other axis b in increasing order, half
// the engineering shear of the cell just below the node along b; then the mean
// over the x = 0 face is removed.  A node beyond the last cell along an axis
// reads the last cell on that axis.
inline std::vector<mgkt::Vector3d> reconstruct_displacement_element_centered(int nx, int ny, int nz,
                                                                             const std::vector<mgkt::Matrix3d>& eps,
                                                                             double pixel) {
    const int n[3] = {nx, ny, nz};
    const int P[3] = {nx + 1, ny + 1, nz + 1};
    const long npts = (long)P[0] * P[1] * P[2];
    std::vector<mgkt::Vector3d> u(npts, mgkt::Vector3d::Zero());
    auto node = [&](const int* q) { return ((long)q[0] * P[1] + q[1]) * P[2] + q[2]; };
    // the cell under node q, one step back along axis `back` (or none)
    auto cell = [&](const int* q, int back) {
        int c[3];
        for (int a = 0; a < 3; ++a) c[a] = (a == back) ? q[a] - 1 : std::min(q[a], n[a] - 1);
        return ((long)c[0] * ny + c[1]) * nz + c[2];
    };
    int q[3];
    // normal strains: sequential integration along each axis
    for (int a = 0; a < 3; ++a)
        for (q[0] = 0; q[0] < P[0]; ++q[0])
            for (q[1] = 0; q[1] < P[1]; ++q[1])
                for (q[2] = 0; q[2] < P[2]; ++q[2]) {
                    if (q[a] == 0) continue;
                    int prev[3] = {q[0], q[1], q[2]};
                    --prev[a];
                    u[node(q)](a) = u[node(prev)](a) + eps[cell(q, a)](a, a) * pixel;
                }
    // shear: gamma/2 of the cell below along each other axis
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            if (b == a) continue;
            const int i = std::min(a, b), j = std::max(a, b);
            for (q[0] = 0; q[0] < P[0]; ++q[0])
                for (q[1] = 0; q[1] < P[1]; ++q[1])
                    for (q[2] = 0; q[2] < P[2]; ++q[2])
                        if (q[b] >= 1) u[node(q)](a) += 0.5 * eps[cell(q, b)](i, j) * pixel;
        }
    // anchor: the mean of the x = 0 face
    mgkt::Vector3d u0 = mgkt::Vector3d::Zero();
    for (int j = 0; j < P[1]; ++j)
        for (int k = 0; k < P[2]; ++k) u0 += u[(long)j * P[2] + k];
    u0 /= (double)(P[1] * P[2]);
    for (auto& v : u) v -= u0;
    return u;
}

namespace mgk_vtu {

inline double von_mises(const mgkt::Matrix3d& t, double factor) {
    const mgkt::Matrix3d dev = t - mgkt::Matrix3d::Identity() * (t.trace() / 3.0);
    return std::sqrt(factor * dev.squaredNorm());
}

// One cell array: ten values a line (vtu_writer.hpp:206-212).
inline void cell_array(std::ofstream& fs, const char* name, long ncell, const std::vector<double>& v) {
    fs << "        <DataArray type=\"Float64\" Name=\"" << name << "\" format=\"ascii\">\n";
    for (long i = 0; i < ncell; ++i) fs << v[i] << ((i % 10 == 9 || i == ncell - 1) ? "\n" : " ");
    fs << "        </DataArray>\n";
}

// The six components and the von Mises measure of a tensor field.
inline void tensor_arrays(std::ofstream& fs, const std::string& prefix, long ncell,
                          const std::vector<mgkt::Matrix3d>& t, double vm_factor) {
    static const char* tags[6] = {"xx", "yy", "zz", "xy", "xz", "yz"};
    static const int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
    std::vector<double> v(ncell);
    for (int c = 0; c < 6; ++c) {
        for (long i = 0; i < ncell; ++i) v[i] = t[i](ii[c], jj[c]);
        cell_array(fs, (prefix + tags[c]).c_str(), ncell, v);
    }
    for (long i = 0; i < ncell; ++i) v[i] = von_mises(t[i], vm_factor);
    cell_array(fs, (prefix + "vm").c_str(), ncell, v);
}

}  // namespace mgk_vtu

// ASCII VTU of a 3D grid (vtu_writer.hpp:169-357).  Optional fields are
// written when non-empty.
inline void export_to_vtu(const std::string& filename, int nx, int ny, int nz, const std::vector<mgkt::Matrix3d>& eps,
                          const std::vector<mgkt::Matrix3d>& sig, const std::vector<double>& E_field,
                          const std::vector<double>& nu_field, double pixel, const std::vector<double>& Tlocal = {},
                          const std::vector<mgkt::Matrix3d>& eps_plastic = {},
                          const std::vector<mgkt::Vector2d>& soft_prop = {}) {
    std::ofstream fs(filename);
    if (!fs.is_open()) throw std::runtime_error("Failed to open VTU file for writing: " + filename);
    const long npts = (long)(nx + 1) * (ny + 1) * (nz + 1);
    const long ncell = (long)nx * ny * nz;
    fs << "<?xml version=\"1.0\"?>\n"
       << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
       << "  <UnstructuredGrid>\n"
       << "    <Piece NumberOfPoints=\"" << npts << "\" NumberOfCells=\"" << ncell << "\">\n";
    const std::vector<mgkt::Vector3d> u = reconstruct_displacement_element_centered(nx, ny, nz, eps, pixel);
    fs << "      <PointData Vectors=\"displacement\">\n"
       << "        <DataArray type=\"Float64\" Name=\"displacement\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (long i = 0; i < npts; ++i) fs << u[i](0) << " " << u[i](1) << " " << u[i](2) << "\n";
    fs << "        </DataArray>\n"
       << "      </PointData>\n"
       << "      <CellData>\n";
    mgk_vtu::cell_array(fs, "E", ncell, E_field);
    mgk_vtu::cell_array(fs, "nu", ncell, nu_field);
    mgk_vtu::tensor_arrays(fs, "eps_", ncell, eps, 2.0 / 3.0);
    mgk_vtu::tensor_arrays(fs, "sig_", ncell, sig, 1.5);
    if (!Tlocal.empty()) mgk_vtu::cell_array(fs, "temperature", ncell, Tlocal);
    if (!eps_plastic.empty()) mgk_vtu::tensor_arrays(fs, "eps_plastic_", ncell, eps_plastic, 2.0 / 3.0);
    if (!soft_prop.empty()) {
        std::vector<double> v(ncell);
        for (int k = 0; k < 2; ++k) {
            for (long i = 0; i < ncell; ++i) v[i] = soft_prop[i](k);
            mgk_vtu::cell_array(fs, k == 0 ? "g_p" : "g_t", ncell, v);
        }
    }
    fs << "      </CellData>\n"
       << "      <Points>\n"
       << "        <DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j)
            for (int k = 0; k <= nz; ++k) fs << i * pixel << " " << j * pixel << " " << k * pixel << "\n";
    fs << "        </DataArray>\n"
       << "      </Points>\n"
       << "      <Cells>\n"
       << "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
    const long sx = (long)(ny + 1) * (nz + 1), sy = nz + 1;   // node strides along x and y
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const long n0 = i * sx + j * sy + k;
                // VTK_HEXAHEDRON order: the x = i face counter-clockwise, then x = i + 1
                const long face[4] = {n0, n0 + 1, n0 + sy + 1, n0 + sy};
                fs << face[0] << " " << face[1] << " " << face[2] << " " << face[3] << " " << face[0] + sx << " "
                   << face[1] + sx << " " << face[2] + sx << " " << face[3] + sx << "\n";
            }
    fs << "        </DataArray>\n"
       << "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
    for (long i = 1; i <= ncell; ++i) fs << i * 8 << ((i % 10 == 0 || i == ncell) ? "\n" : " ");
    fs << "        </DataArray>\n"
       << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (long i = 1; i <= ncell; ++i) fs << 12 << ((i % 10 == 0 || i == ncell) ? "\n" : " ");
    fs << "        </DataArray>\n"
       << "      </Cells>\n"
       << "    </Piece>\n"
       << "  </UnstructuredGrid>\n"
       << "</VTKFile>\n";
}

// 2D fields as one layer of 3D cells (vtu_writer.hpp:359-393).
inline void export_to_vtu_2d(const std::string& filename, int nx, int ny, const std::vector<mgkt::Matrix2d>& eps_2d,
                             const std::vector<mgkt::Matrix2d>& sig_2d, const std::vector<double>& E_field,
                             const std::vector<double>& nu_field, double pixel, const std::vector<double>& Tlocal = {},
                             const std::vector<mgkt::Matrix2d>& eps_plastic_2d = {},
                             const std::vector<mgkt::Vector2d>& soft_prop = {}) {
    const long N = (long)nx * ny;
    auto lift = [N](const std::vector<mgkt::Matrix2d>& t) {
        std::vector<mgkt::Matrix3d> o(N, mgkt::Matrix3d::Zero());
        for (long i = 0; i < N; ++i)
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) o[i](a, b) = t[i](a, b);
        return o;
    };
    const std::vector<mgkt::Matrix3d> ep = eps_plastic_2d.empty() ? std::vector<mgkt::Matrix3d>() : lift(eps_plastic_2d);
    export_to_vtu(filename, nx, ny, 1, lift(eps_2d), lift(sig_2d), E_field, nu_field, pixel, Tlocal, ep, soft_prop);
}
