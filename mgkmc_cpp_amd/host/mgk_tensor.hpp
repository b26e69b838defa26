// mgk_tensor.hpp -- fixed-size 2x2 / 3x3 FP64 tensors for the drop-in host
// API.  Column-major storage with the same memory layout as
// Eigen::Matrix2d / Matrix3d / Vector2d (the reference's member types,
// kmc_solver.hpp:136-149), so vectors of either can be handed to the C ABI
// through data().  Only the operations the reference's KMC and solver code
// use are provided (Zero, Identity, (i,j), trace, squaredNorm, norm,
// transpose, + - * /).
#pragma once
#include <cmath>

namespace mgkt {

template <int R, int C> struct Matrix {
    double a[R * C];
    double& operator()(int i, int j) { return a[i + R * j]; }
    const double& operator()(int i, int j) const { return a[i + R * j]; }
    double& operator()(int i) { return a[i]; }
    const double& operator()(int i) const { return a[i]; }
    double* data() { return a; }
    const double* data() const { return a; }
    static Matrix Zero() {
        Matrix m;
        for (int k = 0; k < R * C; ++k) m.a[k] = 0.0;
        return m;
    }
    static Matrix Identity() {
        Matrix m = Zero();
        for (int k = 0; k < (R < C ? R : C); ++k) m(k, k) = 1.0;
        return m;
    }
    double trace() const {
        double t = 0.0;
        for (int k = 0; k < (R < C ? R : C); ++k) t += (*this)(k, k);
        return t;
    }
    double squaredNorm() const {
        double s = 0.0;
        for (int k = 0; k < R * C; ++k) s += a[k] * a[k];
        return s;
    }
    double norm() const { return std::sqrt(squaredNorm()); }
    Matrix<C, R> transpose() const {
        Matrix<C, R> t;
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < C; ++j) t(j, i) = (*this)(i, j);
        return t;
    }
    Matrix operator+(const Matrix& o) const { Matrix m; for (int k = 0; k < R * C; ++k) m.a[k] = a[k] + o.a[k]; return m; }
    Matrix operator-(const Matrix& o) const { Matrix m; for (int k = 0; k < R * C; ++k) m.a[k] = a[k] - o.a[k]; return m; }
    Matrix& operator+=(const Matrix& o) { for (int k = 0; k < R * C; ++k) a[k] += o.a[k]; return *this; }
    Matrix& operator-=(const Matrix& o) { for (int k = 0; k < R * C; ++k) a[k] -= o.a[k]; return *this; }
    Matrix operator*(double s) const { Matrix m; for (int k = 0; k < R * C; ++k) m.a[k] = a[k] * s; return m; }
    Matrix operator/(double s) const { Matrix m; for (int k = 0; k < R * C; ++k) m.a[k] = a[k] / s; return m; }
    Matrix& operator/=(double s) { for (int k = 0; k < R * C; ++k) a[k] /= s; return *this; }
    friend Matrix operator*(double s, const Matrix& m) { return m * s; }
};

using Matrix2d = Matrix<2, 2>;
using Matrix3d = Matrix<3, 3>;
using Vector2d = Matrix<2, 1>;
using Vector3d = Matrix<3, 1>;

static_assert(sizeof(Matrix2d) == 4 * sizeof(double), "layout");
static_assert(sizeof(Matrix3d) == 9 * sizeof(double), "layout");
static_assert(sizeof(Vector2d) == 2 * sizeof(double), "layout");
static_assert(sizeof(Vector3d) == 3 * sizeof(double), "layout");

}  // namespace mgkt
