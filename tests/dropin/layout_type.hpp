// A second column-major fixed-size tensor type, layout-identical to
// Eigen::Matrix<double, R, C> and to the drop-in's built-in one but written
// independently (its storage is named differently), with the part of Eigen's
// interface the reference's callers and the drop-in headers use.  A caller
// that defines MGK_TENSOR_TEMPLATE to ::layout::Mat gets the drop-in API on
// this type; the test checks that such a caller compiles and runs unchanged.
#pragma once
#include <cmath>

namespace layout {

template <int R, int C> class Mat {
public:
    static Mat Zero() { Mat m; for (int k = 0; k < R * C; ++k) m.v_[k] = 0.0; return m; }
    static Mat Identity() { Mat m = Zero(); for (int k = 0; k < (R < C ? R : C); ++k) m.v_[k + R * k] = 1.0; return m; }
    double& operator()(int i, int j) { return v_[i + R * j]; }
    double operator()(int i, int j) const { return v_[i + R * j]; }
    double& operator()(int i) { return v_[i]; }
    double operator()(int i) const { return v_[i]; }
    double* data() { return v_; }
    const double* data() const { return v_; }
    double trace() const { double t = 0.0; for (int k = 0; k < (R < C ? R : C); ++k) t += v_[k + R * k]; return t; }
    double squaredNorm() const { double s = 0.0; for (double x : v_) s += x * x; return s; }
    double norm() const { return std::sqrt(squaredNorm()); }
    Mat<C, R> transpose() const {
        Mat<C, R> t;
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < C; ++j) t(j, i) = (*this)(i, j);
        return t;
    }
    Mat operator+(const Mat& o) const { Mat m; for (int k = 0; k < R * C; ++k) m.v_[k] = v_[k] + o.v_[k]; return m; }
    Mat operator-(const Mat& o) const { Mat m; for (int k = 0; k < R * C; ++k) m.v_[k] = v_[k] - o.v_[k]; return m; }
    Mat& operator+=(const Mat& o) { for (int k = 0; k < R * C; ++k) v_[k] += o.v_[k]; return *this; }
    Mat& operator-=(const Mat& o) { for (int k = 0; k < R * C; ++k) v_[k] -= o.v_[k]; return *this; }
    Mat operator*(double s) const { Mat m; for (int k = 0; k < R * C; ++k) m.v_[k] = v_[k] * s; return m; }
    Mat operator/(double s) const { Mat m; for (int k = 0; k < R * C; ++k) m.v_[k] = v_[k] / s; return m; }
    Mat& operator/=(double s) { for (double& x : v_) x /= s; return *this; }
    friend Mat operator*(double s, const Mat& m) { return m * s; }

private:
    double v_[R * C];
};

}  // namespace layout
