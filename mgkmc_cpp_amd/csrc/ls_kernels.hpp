// ls_kernels.hpp -- launch interface of the Lippmann-Schwinger kernels.
#pragma once
#include "common.hpp"

namespace mgk {

enum { LS_FULL = 0, LS_DELTA = 1 };

// Device pointers and scalars of one solve.  Real fields are SoA
// [component][voxel] over the symmetric components (common.hpp order);
// spectra are [component][f][m][l-bin] interleaved complex.
struct LsParams {
    Geom g;
    double* eps;            // FP64 strain (in/out)
    const double* epsp;     // FP64 plastic strain or null
    const double* lam;      // per-voxel Lame constants
    const double* mu;
    double* sig;            // FP64 stress (final pass)
    double lam0, mu0;       // reference medium (mean E, mean nu)
    void* spec64;           // FP64 half spectrum, C*NS complex
    void* spec32;           // FP32 half spectrum (may alias spec64's memory)
    float* r32;             // FP32 copy of the last FP64 increment, or null
    float* y32;             // FP32 inner iterate
    float* corr;            // FP32 accumulated correction, or null
    const float* dlam32;    // lam - lam0, mu - mu0 in FP32
    const float* dmu32;
    const void* tw_f64; const void* tw_m64; const void* tw_l64;   // twiddle tables per axis
    const void* tw_f32; const void* tw_m32; const void* tw_l32;
    Ctl* ctl;
    double* partials;
};

bool ls_length_supported(int n);
int ls_partial_slots(const Geom& g);
void ls_enqueue_start(const LsParams& p, bool warm, hipStream_t s);
void ls_enqueue_iteration(const LsParams& p, int mode, hipStream_t s);
void ls_enqueue_final(const LsParams& p, hipStream_t s);
// one pass of the pipeline alone (0 A, 1 B, 2 C, 3 D), for kernel timing
void ls_enqueue_pass(const LsParams& p, int pass, int mode, hipStream_t s);

}  // namespace mgk
