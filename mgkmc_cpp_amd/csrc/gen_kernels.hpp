// gen_kernels.hpp -- device input generation (barrier and STZ-strain draws,
// correlated fields); see gen_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace mgk {

// BarrierGenerator parameters (kmc_solver.hpp:24-33) and the STZ mode.
struct DrawParams {
    int barrier_type;   // 0 gaussian, 1 rayleigh, 2 modified_rayleigh, 3 mixture, other: 0 then clipped
    double b_mean, b_std, b_min_cutoff, b_max_cutoff, b_loc, b_scale, b_epsilon, b_ratio;
    double gamma0;
    int glass_3d;       // 3D: stz_mode_glass_3d instead of the planar mode
};

// generate_correlated_field_cpp's targets (material_generator.hpp:130-137).
struct FieldParams {
    double mean, std, corr;
    int has_min, has_max;
    double clip_min, clip_max;
};

// Rejection rounds of modified_rayleigh before giving up (each round accepts
// with probability about 0.2, so 1000 rounds fail with probability < 1e-90).
constexpr int kMaxRejections = 1000;

// Q0[N*M] and, when cat is not null, the per-(voxel, mode) catalog [C][N*M].
void gen_enqueue_modes(double* Q0, double* cat, long nmodes, int C, const DrawParams& p, unsigned long long seed,
                       hipStream_t s);

// The correlated field of an n[0] x .. x n[dims-1] grid into host `out`;
// noise: N host normals, or null for device normals from `seed`.
void gen_correlated_field(const int* n, int dims, const FieldParams& f, unsigned long long seed, const double* noise,
                          double* out);

// Limits of the variant without allocation (its scratch array is about 16 MiB).
constexpr long kNoAllocMaxVoxels = 1L << 21;   // 128^3
constexpr int kNoAllocMaxAxis = 2048;

// The same field enqueued on stream s with no allocation: noise (N device
// normals, or null for device normals from `seed`) and out (N device values)
// are the caller's device buffers; every other buffer is in the file-scope
// scratch array.  Throws ArgError beyond the limits above.
void gen_correlated_field_noalloc(const int* n, int dims, const FieldParams& f, unsigned long long seed,
                                  const double* noise, double* out, hipStream_t s);

}  // namespace mgk
