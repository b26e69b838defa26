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
    int n_unstable;        // Q < stability_threshold
    int n_included;        // rated modes
    int sel;               // selected flat mode index, -1 if none
    int ambiguous;         // 1: the parallel prefix cannot decide bit-exactly
    unsigned ticket;
    int nblocks;
    int pad0, pad1;
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
};

constexpr int kRateBlock = 1024;   // modes per rate block (256 lanes x 4)

int kmc_nblocks(long nmodes);
void kmc_enqueue_rates(const KmcParams& p, double time, hipStream_t s);
void kmc_enqueue_select(const KmcParams& p, double u2, hipStream_t s);
void kmc_enqueue_select_fallback(const KmcParams& p, int only_if_ambiguous, hipStream_t s);
void kmc_enqueue_flip(const KmcParams& p, int flat, double time, int first_mode, int n_modes, int has_cat,
                      int has_q0, hipStream_t s);
void kmc_enqueue_unstable(const KmcParams& p, hipStream_t s);

}  // namespace mgk
