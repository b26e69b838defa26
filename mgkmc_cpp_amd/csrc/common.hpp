// common.hpp -- shared types of the device library (geometry, control block,
// error handling).  Host and device code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <type_traits>

namespace mgk {

// Grid geometry of one trajectory.  The FFT axes are named by role:
// F (first, x), M (middle, y in 3D, absent in 2D) and L (last, contiguous:
// z in 3D, y in 2D), matching the reference's voxel order idx = x*ny + y and
// idx = x*ny*nz + y*nz + z (linear_elastic_solver.hpp:81, :357).
// A 3D grid with one axis of length 1 keeps that memory order but puts the
// length-1 axis at M, so the transform runs over the two others (the M
// passes are then left out): nz = 1 gives (F, M, L) = (x, z, y), nx = 1
// gives (y, x, z), ny = 1 keeps (x, y, z).
enum GeomPerm { PERM_XYZ = 0, PERM_XZY = 1, PERM_YXZ = 2 };

struct Geom {
    int dims = 2;
    int nf = 1, nm = 1, nl = 1;
    int nlh = 1;       // nl/2 + 1 bins of the real-to-complex last axis
    long N = 1;        // voxels
    long NS = 1;       // complex bins of the half spectrum = nf*nm*nlh
    int C = 3;         // symmetric components stored: 3 (2D) or 6 (3D)
    int nx = 1, ny = 1, nz = 1;   // the grid as given
    int perm = PERM_XYZ;          // physical direction of each FFT axis
    __host__ __device__ long nlines() const { return (long)nf * nm; }
};

// Geometry of an (nx, ny, nz) grid (nz ignored in 2D); no size checks.
inline Geom make_geom(int dims, int nx, int ny, int nz) {
    Geom g;
    g.dims = dims;
    g.nx = nx;
    g.ny = ny;
    g.nz = (dims == 2) ? 1 : nz;
    if (dims == 2) { g.nf = nx; g.nm = 1; g.nl = ny; }
    else if (nz == 1) { g.nf = nx; g.nm = 1; g.nl = ny; g.perm = PERM_XZY; }
    else if (nx == 1) { g.nf = ny; g.nm = 1; g.nl = nz; g.perm = PERM_YXZ; }
    else { g.nf = nx; g.nm = ny; g.nl = nz; }
    g.nlh = g.nl / 2 + 1;
    g.N = (long)g.nf * g.nm * g.nl;
    g.NS = (long)g.nf * g.nm * g.nlh;
    g.C = (dims == 2) ? 3 : 6;
    return g;
}

// Every buffer a kernel reads or writes through its parameter block is
// device (global) memory.  A pointer the compiler cannot trace to a kernel
// argument -- a batched launch reads its parameter block from memory -- would
// otherwise be accessed as a flat address, which counts on both the vector
// memory and the LDS counters, so every LDS wait of the block also waits for
// its outstanding HBM accesses.  These helpers access global memory directly.
// Scalars and vector types are accessed as they are; a small struct (a complex
// value) as the unsigned vector of its size, naturally aligned in its array.
template <typename R> struct cplx;
template <typename T> using gmem = __attribute__((address_space(1))) T;
typedef unsigned gwords2 __attribute__((ext_vector_type(2)));
typedef unsigned gwords4 __attribute__((ext_vector_type(4)));
template <typename T> struct GWords { using type = T; };
template <typename R> struct GWords<cplx<R>> { using type = std::conditional_t<sizeof(R) == 8, gwords4, gwords2>; };
template <typename T> __device__ __forceinline__ T gld(const T* p) {
    using W = typename GWords<T>::type;
    return __builtin_bit_cast(T, *(const gmem<W>*)p);
}
template <typename T> __device__ __forceinline__ void gst(T* p, const T& v) {
    using W = typename GWords<T>::type;
    *(gmem<W>*)p = __builtin_bit_cast(W, v);
}

// Symmetric component order on the device.
// 2D: 0 xx, 1 yy, 2 xy.   3D: 0 xx, 1 yy, 2 zz, 3 xy, 4 xz, 5 yz.
__host__ __device__ inline int comp_of(int dims, int i, int j) {
    if (i == j) return i;
    int a = i < j ? i : j, b = i < j ? j : i;
    if (dims == 2) return 2;
    return (a == 0 && b == 1) ? 3 : (a == 0 ? 4 : 5);
}

// Control block of one Lippmann-Schwinger solve, in device memory and
// mirrored into pinned host memory at the end of every batch of iterations.
struct Ctl {
    int done;            // convergence reached or iteration cap hit
    int iters;           // FP64 iterations run
    int inner;           // FP32 inner iterations run (refinement path)
    int corr_valid;      // the FP32 correction field holds a pending update
    unsigned ticket;     // last-block ticket of the reduction
    int max_iter;
    int keep_shift;      // the warm-start shift is already set: a mean pass leaves it alone
    int skip_final;      // an event attempt without an event: the fields did not change
    double tol;
    double diff;         // last relative increment
    double norm_extra;   // sum over the uniform antisymmetric part of eps (host-known)
    double shift[6];     // eps_bar - mean(eps): the warm-start shift of iteration 0
    double eps_bar[6];   // imposed mean strain (symmetric part)
    double mean_eps[6];  // outputs of the final pass
    double mean_sig[6];
    double diff_sq, norm_sq;
    double diff_hist[32];  // relative increment of each iteration (FP64 and CG)
    double eps_norm;       // ||eps|| after the last FP64 iteration
    double cg_alpha, cg_beta, cg_gamma;
    int cg_stop;           // CG breakdown: the rest of the cycle is skipped
    int cycle_len;         // CG steps of this solve's cycles (0: as many as enqueued)
    // CG scalars and step count by step parity, for reductions folded into
    // the next step's first kernel: the reduction of step j writes slot j & 1
    // and reads slot (j - 1) & 1, which no block of that kernel writes
    double cgs_gamma[2], cgs_alpha[2], cgs_beta[2];
    int cgs_inner[2];
    int cycle_len2;        // CG steps of the later cycles (after the second FP64 iteration; 0: as enqueued)
    int cycle_start;       // inner count at the start of the current cycle
    int cyc_steps[2];      // most CG steps of a finished first cycle, of a finished later cycle
    int cyc_exhausted[2];  // a first / later cycle ran out of steps without a decision
    int cg_err;            // a persistent CG cycle's grid barrier timed out
    double cycle_r0;       // relative FP64 increment that started the current refinement cycle
};

struct Status {
    int code = 0;
    std::string msg;
};

}  // namespace mgk

#define MGK_HIP_CHECK(expr)                                                                 \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            throw mgk::HipError(std::string(#expr) + ": " + hipGetErrorString(e_));         \
        }                                                                                   \
    } while (0)

namespace mgk {
struct HipError {
    std::string what;
    explicit HipError(std::string w) : what(std::move(w)) {}
};
struct ArgError {
    std::string what;
    explicit ArgError(std::string w) : what(std::move(w)) {}
};
struct UnsupportedError {
    std::string what;
    explicit UnsupportedError(std::string w) : what(std::move(w)) {}
};
}  // namespace mgk
