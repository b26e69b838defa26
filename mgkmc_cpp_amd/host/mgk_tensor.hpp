// mgk_tensor.hpp -- the 2x2 / 3x3 FP64 tensor types of the drop-in host API.
//
// The reference's members and free functions take Eigen::Matrix2d, Matrix3d and
// Vector2d (kmc_solver.hpp:136-149, linear_elastic_solver.hpp:146-160).  The
// drop-in headers name them mgkt::Matrix2d, mgkt::Matrix3d, mgkt::Vector2d,
// which are:
//   * Eigen's own types when Eigen is included before the drop-in headers
//     (EIGEN_WORLD_VERSION defined): a reference caller's
//     std::vector<Eigen::Matrix3d> is passed as is;
//   * MGK_TENSOR_TEMPLATE<R, C> when a caller defines that macro to a
//     column-major template of its own with the same interface;
//   * otherwise the small built-in class below.
// Whatever the choice, the layout is Eigen's: R*C contiguous doubles in
// column-major order, so vectors of them go to the C ABI through data().
// The drop-in code uses only this part of Eigen's interface: Zero, Identity,
// (i, j), (i), data, trace, squaredNorm, norm, transpose, + - += -=, scalar
// * and /, /=.
#pragma once
#include <cmath>

#if defined(MGK_TENSOR_TEMPLATE)

namespace mgkt {
template <int R, int C> using Matrix = MGK_TENSOR_TEMPLATE<R, C>;
}

#elif defined(EIGEN_WORLD_VERSION)

namespace mgkt {
template <int R, int C> using Matrix = Eigen::Matrix<double, R, C>;
}

#else

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

}  // namespace mgkt

#endif

namespace mgkt {

using Matrix2d = Matrix<2, 2>;
using Matrix3d = Matrix<3, 3>;
using Vector2d = Matrix<2, 1>;
using Vector3d = Matrix<3, 1>;

static_assert(sizeof(Matrix2d) == 4 * sizeof(double), "Matrix2d: 4 contiguous doubles");
static_assert(sizeof(Matrix3d) == 9 * sizeof(double), "Matrix3d: 9 contiguous doubles");
static_assert(sizeof(Vector2d) == 2 * sizeof(double), "Vector2d: 2 contiguous doubles");
static_assert(sizeof(Vector3d) == 3 * sizeof(double), "Vector3d: 3 contiguous doubles");

}  // namespace mgkt
