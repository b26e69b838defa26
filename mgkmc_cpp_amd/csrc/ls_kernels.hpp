// ls_kernels.hpp -- launch interface of the Lippmann-Schwinger kernels.
#pragma once
#include "common.hpp"

namespace mgk {

enum { LS_FULL = 0, LS_CG = 1 };

// per-launch flags of passes A and E
enum {
    LS_FLAG_SHIFT = 1,        // FULL: add the warm-start shift (iteration 0 of a warm solve)
    LS_FLAG_CORR = 2,         // FULL: add the pending FP32 correction (refinement path)
    LS_FLAG_CG_FIRST = 4,     // CG step 0 of a cycle: operator on r0 only
    LS_FLAG_CG_UPDATE1 = 8,   // CG step 1: p = r, s = w, x = alpha p (no history)
    LS_FLAG_CG_UPDATE = 16,   // CG step >= 2: full recurrence
    LS_FLAG_SPEC = 32,        // CG step past the likely count: test `done` before loading anything
    LS_FLAG_ODD = 64,         // CG step index within the cycle is odd (the reduction slots)
    LS_FLAG_LAST = 128,       // last CG step of the cycle: its reduction is a kernel of its own
    LS_FLAG_PROBE = 256,      // timing launch: a folded reduction is loaded and computed, not applied
    LS_FLAG_PLANE_SLOTS = 512 // set by the launcher: the CG reduction uses the parity slots (plane passes)
};
// a CG step's index within its cycle rides in the flags' high bits
constexpr int LS_STEP_SHIFT = 16;

// how a solve starts
enum {
    LS_START_KEEP = 0,        // warm start, shift already in the control block
    LS_START_MEAN = 1,        // warm start, compute the field mean and the shift
    LS_START_FILL = 2,        // cold start: eps = eps_bar everywhere
    LS_START_APPLY = 3        // apply the shift to eps now (warm start with no iteration)
};

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
    float* r32;             // CG residual r (the FP64 increment starts it), or null
    float* w32;             // CG w = A r
    float* p32;             // CG direction p
    float* s32;             // CG s = A p
    int mandel;             // the CG vectors (r, w, p, s, x) hold Mandel components (plane passes)
    float* corr;            // CG iterate x: FP32 correction of eps, or null
    const float* dlam32;    // lam - lam0, mu - mu0 in FP32
    const float* dmu32;
    const void* tw_f64; const void* tw_m64; const void* tw_l64;   // twiddle tables per axis
    const void* tw_f32; const void* tw_m32; const void* tw_l32;
    Ctl* ctl;
    double* partials;
    unsigned* gbar;         // grid barrier words of the persistent CG cycle (inside the partials buffer), or null
    int r0_cart;            // FP64 plane passes: a cycle's starting increment is Cartesian, in p32's storage
};

// One launch of a pipeline step over B trajectories of one shape: row
// blockIdx.y of every grid works on the trajectory whose parameter block is
// P[blockIdx.y] (device memory).  A single context (batch = false, B = 1)
// passes its block by value instead (kernel arguments are preloaded).
// Block-shape variants of a shape's passes, chosen by the autotuner
// (index 0: FP64 iterations, 1: FP32 CG steps); 0 is the default shape.
struct LsTune {
    int fwd[2] = {0, 0};   // pass A: 1 = half the line pairs per block
    int mid[2] = {0, 0};   // passes B, D: 1 = half the columns per block
};

struct LsLaunch {
    Geom g;
    const LsParams* P;
    int B;
    LsParams h;   // host copy of a single context's parameter block (passed by value)
    bool batch;   // a batch (any B, also 1): parameter blocks from P
    bool plane = false;   // CG steps run the plane passes (ls_plane_path)
    bool plane64 = false; // FP64 iterations run the FP64 plane passes (ls_plane64_path)
    LsTune tune{};        // block-shape variants of this shape (autotuner)
};

bool ls_length_supported(int n);
// 3D grids whose CG steps run as plane passes (square nm = nl planes of a
// supported length; the environment variable MGK_LS_PLANE=0 turns them off)
bool ls_plane_supported(const Geom& g);
bool ls_plane_path(const Geom& g);
// 3D grids whose FP64 iterations run as plane passes: the plane grids of up
// to 64-point planes whose FP64 pass C tiles fit (MGK_LS_PLANE64=0 turns them off)
bool ls_plane64_path(const Geom& g);
int ls_partial_slots(const Geom& g);
// offset (doubles) of the grid barrier words inside the partials buffer
long ls_gbar_offset(const Geom& g);
// Persistent CG cycles (k_cg_cycle): a single context on a cubic plane grid
// whose 6 n blocks the device holds at once (MGK_LS_CYCLE=0 turns them off).
bool ls_cycle_supported(const Geom& g);
// All CG steps of one refinement cycle (at most J, as J enqueued plane-pass
// steps would run) in one launch whose blocks are all resident; the caller
// runs at most one cycle at a time in the process.
void ls_enqueue_cycle(const LsLaunch& l, int J, hipStream_t s);
void ls_enqueue_start(const LsLaunch& l, int start, hipStream_t s);
void ls_enqueue_iteration(const LsLaunch& l, int mode, int flags, hipStream_t s);
void ls_enqueue_final(const LsLaunch& l, hipStream_t s);
// one pass of the pipeline alone, for kernel timing: 0 A, 1 B, 2 C, 3 D,
// 4 E (without its reduction), 5 A with the CG vector update (FP32); plane
// passes (FP32): 6 forward plane with the CG update, 7 C on the plane
// layout, 8 inverse plane (without its reduction); FP64 plane passes: 9
// forward plane, 10 C on the tiled FP64 spectrum, 11 inverse plane
void ls_enqueue_pass(const LsLaunch& l, int pass, int mode, hipStream_t s);

}  // namespace mgk
