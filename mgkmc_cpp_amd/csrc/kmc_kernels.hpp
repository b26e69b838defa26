// kmc_kernels.hpp -- launch interface of the KMC event kernels.
#pragma once
#include "common.hpp"

namespace mgk {

// Control block of the rate/selection kernels (device, mirrored to pinned host).
struct KmcCtl {
    double total_rate;     // FP64 sum of included rates (block tree order)
    double u2;             // selection uniform
    double target;         // u2 * total_rate used by the selection
    double q_sel;
    double t_wait;         // residence time of a device-decided event
    double time_new;       // time + t_wait
    int n_unstable;        // Q < stability_threshold
    int n_included;        // rated modes
    int sel;               // selected flat mode index, -1 if none
    int ambiguous;         // 1: the parallel prefix cannot decide bit-exactly
    unsigned ticket;
    int nblocks;
    int go;                // 1: selection and flip proceed (0: no event this attempt)
    int outcome;           // MGK_EVENT_* of a device-decided event
    double flip_c[6];      // symmetric components of the last flipped strain (before the redraw)
    int flip_voxel;        // its voxel
    int pad_;
};

// The LS and KMC control blocks of one context, adjacent so that one copy
// reads both back.
struct CtlPair {
    Ctl ls;
    KmcCtl kmc;
};
static_assert(sizeof(Ctl) % 8 == 0, "Ctl is a whole number of doubles");

// Fast-patching kernels in device memory (kmc_solver.hpp:276-332): per basis
// b, the stress response K[b][c][w] of a unit deviatoric eigenstrain, on a
// window of (2R+1)^2 or (2R+1)^2 (2R+1 | 1) cells, and the missing mean
// Km[b][c]; sunit[c] is sigma_macro_unit (:867-878).
struct PatchParams {
    const double* K;
    const double* Km;
    const double* sunit;
    int R, side, side_z, zr, nb, W;
};

// Arguments of the device-side event decision (one attempt of the
// reference's inner loop, kmc_solver.hpp:763-952 / :1776-1980).
struct EventArgs {
    double time;           // simulation time before the attempt
    double neg_log_u1;     // -log(u1) of the residence-time draw
    double remaining;      // time left in the strain step
    double strain_rate;    // eps_macro(drv) += strain_rate * t_wait
    double a[9];           // eps_macro before the attempt, D x D column-major
    double n_vox;          // N, for the antisymmetric norm term
    double u2;             // selection uniform
    int drv;               // driven entry of a (row + D * col)
    int cascade_check;     // instability_mode == "cascade": unstable modes pre-empt the event
    int keep_shift;        // the field mean is known: shift = eps_bar - mean (else a mean pass follows)
    int active;            // 0: no attempt for this trajectory (MGK_EVENT_SKIPPED)
    int init_ctl;          // 1: the attempt resets the LS control block itself (max_iter, tol, mean)
    int max_iter;
    int cycle_len;         // CG steps per refinement cycle of this trajectory (0: all enqueued)
    int cycle_len2;        // the same for its later cycles
    double tol;
    double mean_eps[6];    // field mean of the last solve (used when keep_shift)
};

// Redraw values passed by value to the flip (no staging copy): n_modes
// catalog rows of C, then n_modes Q0 values.
constexpr int kRedrawArgs = 112;
struct RedrawArgs {
    double v[kRedrawArgs];
};

// Per-trajectory KMC physics scalars (device copy of mgk_kmc_params plus
// derived values).
struct KmcPhys {
    int M;
    int directional;
    int instability_kmc;
    int pad;
    double cap, jp, jt, nfrac, tau, volume, eff_volume, nu0, beta, thr;
    int tau_inf;   // tau == +inf
    int pad2;
};

struct KmcParams {
    Geom g;
    KmcPhys ph;
    const double* sig;       // [C][N]
    double* eps_p;           // [C][N]
    double* Q;               // [N*M]
    double* Q0;              // [N*M]
    double* cat;             // [C][N*M], or [C][M] when cat_shared
    int cat_shared;          // every voxel holds the same M catalog strains
    double* soft;            // [2][N]  g_p, g_t
    double* last_t;          // [N]
    double* prev_dir;        // [C][N]
    double* bsum;            // [nblocks] block sums of rates
    double* bincl;           // [nblocks] inclusive block prefix
    int* bunst;              // [nblocks] unstable counts
    int* bunst_off;          // [nblocks] exclusive prefix of unstable counts
    int* binc;               // [nblocks] included counts
    KmcCtl* kctl;
    const double* staging;   // redraw values: n_modes*C catalog then n_modes Q0
    int* unst_idx;           // compacted unstable indices
    double* unst_q;
    int skip;                // batch row of an inactive trajectory: the rate pass leaves its state alone
};

constexpr int kRateBlock = 1024;   // modes per rate block
constexpr int kRateThreads = 512;  // lanes per rate block (2 modes each)
constexpr int kStageMinM = 4;      // from this many modes per voxel the rate pass stages voxels in LDS
constexpr int kStageCat = 64;      // shared catalogs of at most this many modes are staged too

// One launch over B trajectories of one shape and mode count: row
// blockIdx.y of every grid works on the trajectory whose parameter block is
// P[blockIdx.y] (device memory).  A single context (batch = false, B = 1)
// passes its block by value instead.
struct KmcLaunch {
    Geom g;
    int M;
    const KmcParams* P;
    int B;
    KmcParams h;   // host copy of a single context's parameter block (passed by value)
    bool batch;    // a batch (any B, also 1): parameters, times and event arguments from device arrays
};

int kmc_nblocks(long nmodes);
// times: per-trajectory times (device, B entries) or null for `time`
// with_scan = false leaves the block-sum scan to kmc_enqueue_event_chain
void kmc_enqueue_rates(const KmcLaunch& l, double time, const double* times, hipStream_t s, bool with_scan = true);
// ea: per-trajectory event arguments (device) whose u2 replaces `u2`, or null
void kmc_enqueue_select(const KmcLaunch& l, double u2, const EventArgs* ea, hipStream_t s);
void kmc_enqueue_select_fallback(const KmcLaunch& l, int only_if_ambiguous, hipStream_t s);
// time_from_ctl: the flip time is the device-decided event's time_new
// host_redraw: redraw values passed by value (single trajectory, at most
// kRedrawArgs doubles) instead of the device staging buffer
void kmc_enqueue_flip(const KmcLaunch& l, int flat, double time, int first_mode, int n_modes, int has_cat,
                      int has_q0, hipStream_t s, int time_from_ctl = 0, const double* host_redraw = nullptr);
bool kmc_redraw_fits_args(int n_modes, int C);
// A cascade round's flips (device flat indices and redraw values, n_modes =
// M (all_modes) or 1 per flip) applied in order by one block.
void kmc_enqueue_flip_list(const KmcLaunch& l, const int* flats, const double* stage, int n, double time,
                           int all_modes, int has_cat, int has_q0, hipStream_t s);
void kmc_enqueue_unstable(const KmcLaunch& l, hipStream_t s);
// Fast-patching stress update: sig += sunit * pred (pred != 0), the kernel
// superposition around voxel for the flipped strain c (symmetric
// components), and the missing-mean shift.  voxel < 0: the predictor only.
// from_ctl: voxel, strain and pred = strain_rate * t_wait (t_wait > 0) from
// the control block of the last device-decided event (nothing if no event).
void kmc_enqueue_patch(const KmcLaunch& l, const PatchParams& pp, int voxel, const double* c, double pred,
                       int from_ctl, double strain_rate, hipStream_t s);
// decision of one event attempt per trajectory after kmc_enqueue_rates; on
// "no event" it marks that trajectory's LS control block done so a solve
// enqueued behind it returns at once.  B = 1: ctl and ea_one; a batch: the
// control blocks P[y].ctl and the arguments ea[y] (device).
struct LsParams;
// scan + event decision + selection (+ flip with its redraw) in one launch
// after kmc_enqueue_rates(..., with_scan = false); host_redraw as for
// kmc_enqueue_flip (single trajectory), else the staging buffer
void kmc_enqueue_event_chain(const KmcLaunch& l, Ctl* ctl, const LsParams* P, const EventArgs& ea_one,
                             const EventArgs* ea, bool flip, int first_mode, int n_modes, int has_cat, int has_q0,
                             const double* host_redraw, hipStream_t s);

}  // namespace mgk
