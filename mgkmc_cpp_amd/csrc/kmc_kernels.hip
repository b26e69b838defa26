// kmc_kernels.hip -- per-event KMC kernels on gfx950: barriers and rates,
// FP64 block sums, two-level event selection, flip, unstable-mode compaction.
//
// Rates span many decades (exp(-Q/kT)), so every sum is FP64.  Selection is
// the first included mode whose sequential prefix sum reaches
// target = u2 * total (kmc_solver.hpp:816-826).  The device computes block
// sums and an in-block scan; when the target lies within a rounding bound of
// a prefix boundary -- the only case where the parallel order could choose a
// different mode than the reference's left-to-right sum -- it reports
// "ambiguous" and the exact sequential pass (k_select_seq) decides instead.
// A device-decided event runs the scan, decision, selection and flip in one
// block (k_event_chain).
#include "common.hpp"
#include "kmc_kernels.hpp"
#include "ls_kernels.hpp"

namespace mgk {

// The parameter block of this block row: the kernel argument for a single
// trajectory (kernel arguments are preloaded), P[blockIdx.y] for a batch.
template <bool BATCH, typename T> __device__ __forceinline__ const T& param_sel(const T& a, const T* P) {
    if constexpr (BATCH) return P[blockIdx.y];
    else return a;
}

template <typename T> __device__ __forceinline__ T wsum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Rate of one mode from its barrier (kmc_solver.hpp:795-801): included when
// instability_mode == "kmc" or Q > 0; V_eff nu0 for Q <= 0, else
// V_eff nu0 exp(-Q beta) (beta = +inf at T = 0 gives 0).
__device__ __forceinline__ double mode_rate(const KmcPhys& ph, double q, bool& incl) {
    incl = ph.instability_kmc || q > 0.0;
    if (!incl) return 0.0;
    return (q <= 0.0) ? ph.eff_volume * ph.nu0 : ph.eff_volume * ph.nu0 * exp(-q * ph.beta);
}

// ---------------------------------------------------------------------------
// update_barriers + rates (kmc_solver.hpp:341-410, :783-806; 3D :1340-1411,
// :1796-1818), split into the voxel part (softening state) and the mode part
// (directional factor, resolved work, barrier), each in the reference's
// operation order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double voxel_gbase(const KmcPhys& ph, double time, double gp, double g_t, double t_last) {
    double g_t_curr = 0.0;
    if (g_t > 0.0) {
        if (ph.tau_inf) g_t_curr = g_t;
        else if (t_last == -INFINITY) g_t_curr = 0.0;
        else {
            double dt = time - t_last;
            if (dt < 0.0) dt = 0.0;
            g_t_curr = g_t * exp(-dt / ph.tau);
        }
    }
    double g_base = gp + g_t_curr;
    if (ph.cap > 0.0 && g_base > ph.cap) g_base = ph.cap;
    return g_base;
}

// Q of one mode.  pd: the voxel's previous direction (read when
// directional); eg: exp(-g_base), used when not directional (then
// g_eff = 1.0 * g_base = g_base exactly).
template <int D, int C, class S, class V, class P>
__device__ __forceinline__ double mode_barrier(const KmcPhys& ph, double g_base, double eg, double q0, const S& sig,
                                               const V& cat, const P& pd) {
    const double GPa_nm3_to_eV = 6.241509;
    double decay = eg;
    if (ph.directional) {
        double modifier = 1.0;
        double dot = 0.0, nmq = 0.0, npq = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                const int c = comp_of(D, a, b);
                const double vm = cat(c), vp = pd(c);
                dot += vm * vp;
                nmq += vm * vm;
                npq += vp * vp;
            }
        const double npv = sqrt(npq), nmv = sqrt(nmq);
        if (!(npv < 1e-12 || nmv < 1e-12)) {
            const double ct = dot / (nmv * npv);
            modifier = ((1.0 + ct) * (1.0 + ct)) / 4.0;
        }
        const double g_eff = modifier * g_base;
        decay = exp(-g_eff);
    }
    double w_sum = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            const int c = comp_of(D, a, b);
            w_sum += sig(c) * cat(c);
        }
    const double w = 0.5 * ph.volume * (w_sum / 1e9) * GPa_nm3_to_eV;
    return q0 * decay - w;
}

// Block totals of rates, unstable and included counts (thread 0 writes).
__device__ __forceinline__ void rate_block_totals(const KmcParams& p, double lsum, int lun, int linc) {
    constexpr int NW = kRateThreads / 64;
    __shared__ double sd[NW];
    __shared__ int su[NW], si[NW];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    lsum = wsum(lsum);
    lun = wsum(lun);
    linc = wsum(linc);
    if (lane == 0) { sd[wid] = lsum; su[wid] = lun; si[wid] = linc; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        int tu = 0, ti = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) { t += sd[w]; tu += su[w]; ti += si[w]; }
        p.bsum[blockIdx.x] = t;
        p.bunst[blockIdx.x] = tu;
        p.binc[blockIdx.x] = ti;
    }
}

// One lane per (voxel, mode), kRateBlock modes per block of kRateThreads
// lanes; every input of a lane's modes is loaded before any is used.  The
// catalog is per (voxel, mode) or, when every voxel holds the same M strains,
// one shared copy of M entries.  For M < kStageMinM.
template <int D, int C, bool BATCH>
__global__ __launch_bounds__(kRateThreads) void k_rates(const KmcParams p_arg, const KmcParams* __restrict__ PK,
                                                        double time, const double* __restrict__ times) {
    const KmcParams& p = param_sel<BATCH>(p_arg, PK);
    if (BATCH && p.skip) return;   // inactive batch row: Q and the block sums stay as they are
    if (BATCH) time = times[blockIdx.y];
    constexpr int K = kRateBlock / kRateThreads;
    const Geom g = p.g;
    const KmcPhys ph = p.ph;
    const long nmodes = g.N * ph.M;
    const long base = (long)blockIdx.x * kRateBlock;

    long mi[K], vi[K];
    bool ok[K];
    double gp[K], gt[K], tl[K], q0[K], s[K][C], cat[K][C], pd[K][C];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const long i = base + k * kRateThreads + threadIdx.x;
        ok[k] = i < nmodes;
        mi[k] = ok[k] ? i : 0;
        vi[k] = (long)((unsigned)mi[k] / (unsigned)ph.M);   // N*M < 2^31 (mgk_kmc_init)
        gp[k] = p.soft[vi[k]];
        gt[k] = p.soft[g.N + vi[k]];
        tl[k] = p.last_t[vi[k]];
        q0[k] = p.Q0[mi[k]];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            s[k][c] = p.sig[c * g.N + vi[k]];
            cat[k][c] = p.cat_shared ? p.cat[c * ph.M + (mi[k] - vi[k] * ph.M)] : p.cat[c * nmodes + mi[k]];
        }
    }
    if (ph.directional) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < C; ++c) pd[k][c] = p.prev_dir[c * g.N + vi[k]];
    }

    double lsum = 0.0;
    int lun = 0, linc = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!ok[k]) continue;
        const double g_base = voxel_gbase(ph, time, gp[k], gt[k], tl[k]);
        const double eg = ph.directional ? 0.0 : exp(-g_base);
        const double q = mode_barrier<D, C>(ph, g_base, eg, q0[k], [&](int c) { return s[k][c]; },
                                            [&](int c) { return cat[k][c]; }, [&](int c) { return pd[k][c]; });
        p.Q[mi[k]] = q;
        bool incl;
        const double r = mode_rate(ph, q, incl);
        if (incl) { lsum += r; linc++; }
        if (q < ph.thr) lun++;
    }
    rate_block_totals(p, lsum, lun, linc);
}

// The same pass for M >= kStageMinM: the block's voxels (at most
// kRateBlock / M + 2) are staged in LDS once -- stress, previous direction,
// softening base and its exponential -- so a lane loads only its modes' Q0
// (and catalog rows when they are per voxel).  Same arithmetic, same bits.
template <int D, int C, bool BATCH>
__global__ __launch_bounds__(kRateThreads) void k_rates_staged(const KmcParams p_arg, const KmcParams* __restrict__ PK,
                                                               double time, const double* __restrict__ times) {
    const KmcParams& p = param_sel<BATCH>(p_arg, PK);
    if (BATCH && p.skip) return;   // inactive batch row: Q and the block sums stay as they are
    if (BATCH) time = times[blockIdx.y];
    constexpr int K = kRateBlock / kRateThreads;
    constexpr int NV = kRateBlock / kStageMinM + 2;
    static_assert(NV <= kRateThreads, "one staged voxel per lane");
    const Geom g = p.g;
    const KmcPhys ph = p.ph;
    const long nmodes = g.N * ph.M;
    const long base = (long)blockIdx.x * kRateBlock;
    __shared__ double s_sig[C][NV], s_pd[C][NV], s_gb[NV], s_eg[NV];
    __shared__ double s_cat[C][kStageCat];

    const unsigned M = (unsigned)ph.M;
    const long v0 = (long)((unsigned)base / M);
    const long last = (base + kRateBlock < nmodes ? base + kRateBlock : nmodes) - 1;
    const long vend = (long)((unsigned)last / M);
    const int nv = (int)(vend - v0 + 1);
    const bool cat_lds = p.cat_shared && ph.M <= kStageCat;   // M*C <= kRateThreads entries
    // per-mode inputs
    long mi[K];
    int vj[K];
    bool ok[K];
    double q0[K], cat[K][C];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const long i = base + k * kRateThreads + threadIdx.x;
        ok[k] = i < nmodes;
        mi[k] = ok[k] ? i : base;
        const long v = (long)((unsigned)mi[k] / M);
        vj[k] = (int)(v - v0);
        q0[k] = p.Q0[mi[k]];
        if (!p.cat_shared) {
#pragma unroll
            for (int c = 0; c < C; ++c) cat[k][c] = p.cat[c * nmodes + mi[k]];
        } else if (!cat_lds) {
#pragma unroll
            for (int c = 0; c < C; ++c) cat[k][c] = p.cat[c * ph.M + (mi[k] - v * ph.M)];
        }
    }
    // per-voxel inputs, one voxel per lane
    const int j = threadIdx.x;
    const long v = v0 + (j < nv ? j : 0);
    double gp = p.soft[v], gt = p.soft[g.N + v], tl = p.last_t[v], sv[C], pv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) sv[c] = p.sig[c * g.N + v];
    if (ph.directional) {
#pragma unroll
        for (int c = 0; c < C; ++c) pv[c] = p.prev_dir[c * g.N + v];
    }
    double cv = 0.0;
    if (cat_lds && j < ph.M * C) cv = p.cat[j];   // [C][M] order
    if (j < nv) {
        const double g_base = voxel_gbase(ph, time, gp, gt, tl);
        s_gb[j] = g_base;
        s_eg[j] = ph.directional ? 0.0 : exp(-g_base);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            s_sig[c][j] = sv[c];
            if (ph.directional) s_pd[c][j] = pv[c];
        }
    }
    if (cat_lds && j < ph.M * C) s_cat[j / ph.M][j % ph.M] = cv;
    __syncthreads();

    double lsum = 0.0;
    int lun = 0, linc = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!ok[k]) continue;
        const int jj = vj[k];
        const int m = (int)(mi[k] - (v0 + jj) * (long)ph.M);
        const double q = cat_lds
            ? mode_barrier<D, C>(ph, s_gb[jj], s_eg[jj], q0[k], [&](int c) { return s_sig[c][jj]; },
                                 [&](int c) { return s_cat[c][m]; }, [&](int c) { return s_pd[c][jj]; })
            : mode_barrier<D, C>(ph, s_gb[jj], s_eg[jj], q0[k], [&](int c) { return s_sig[c][jj]; },
                                 [&](int c) { return cat[k][c]; }, [&](int c) { return s_pd[c][jj]; });
        p.Q[mi[k]] = q;
        bool incl;
        const double r = mode_rate(ph, q, incl);
        if (incl) { lsum += r; linc++; }
        if (q < ph.thr) lun++;
    }
    rate_block_totals(p, lsum, lun, linc);
}

// Inclusive wave scan (lane order) of a double and two ints.
__device__ __forceinline__ void wave_scan(double& v, int& a, int& b) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double vv = __shfl_up(v, off, 64);
        const int aa = __shfl_up(a, off, 64), bb = __shfl_up(b, off, 64);
        if (lane >= off) { v += vv; a += aa; b += bb; }
    }
}

// One block: inclusive prefix of the block sums (chunked, fixed order), the
// exclusive prefix of unstable counts and the totals.  A kernel of its own so
// that the rate blocks need no device-scope fence.
__device__ __forceinline__ void rates_scan_body(const KmcParams& p, int nb) {
    constexpr int G = 8;   // block sums per lane loaded together
    const int chunk = (nb + 1023) / 1024;
    const int b0 = threadIdx.x * chunk;
    double csum = 0.0;
    int cun = 0, cinc = 0;
    double vs0[G];   // the first group's sums and counts, kept for the prefix pass
    int vu0[G];
    for (int g0 = 0; g0 < chunk; g0 += G) {
        double vs[G];
        int vu[G], vi[G];
        bool ok[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int b = b0 + g0 + k;
            ok[k] = (g0 + k < chunk) && (b < nb);
            vs[k] = p.bsum[ok[k] ? b : 0];
            vu[k] = p.bunst[ok[k] ? b : 0];
            vi[k] = p.binc[ok[k] ? b : 0];
        }
        if (g0 == 0) {
#pragma unroll
            for (int k = 0; k < G; ++k) { vs0[k] = vs[k]; vu0[k] = vu[k]; }
        }
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (ok[k]) { csum += vs[k]; cun += vu[k]; cinc += vi[k]; }
    }
    // exclusive prefix of the lane totals: wave scans, then the wave totals
    __shared__ double ws[16];
    __shared__ int wu[16], wi[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double incl_s = csum;
    int incl_u = cun, incl_i = cinc;
    wave_scan(incl_s, incl_u, incl_i);
    if (lane == 63) { ws[wid] = incl_s; wu[wid] = incl_u; wi[wid] = incl_i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        int ru = 0, ri = 0;
        for (int w = 0; w < 16; ++w) {
            const double t = ws[w];
            const int tu = wu[w], ti = wi[w];
            ws[w] = r; wu[w] = ru; wi[w] = ri;   // exclusive wave offsets
            r += t; ru += tu; ri += ti;
        }
    }
    __syncthreads();
    double ex_s = __shfl_up(incl_s, 1, 64);
    int ex_u = __shfl_up(incl_u, 1, 64);
    if (lane == 0) { ex_s = 0.0; ex_u = 0; }
    double run = ws[wid] + ex_s;
    int runu = wu[wid] + ex_u;
    for (int g0 = 0; g0 < chunk; g0 += G) {
        double vs[G];
        int vu[G];
        bool ok[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int b = b0 + g0 + k;
            ok[k] = (g0 + k < chunk) && (b < nb);
            vs[k] = (g0 == 0) ? vs0[k] : p.bsum[ok[k] ? b : 0];
            vu[k] = (g0 == 0) ? vu0[k] : p.bunst[ok[k] ? b : 0];
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (!ok[k]) continue;
            const int b = b0 + g0 + k;
            run += vs[k];
            p.bincl[b] = run;
            p.bunst_off[b] = runu;
            runu += vu[k];
        }
    }
    __shared__ double s_total;
    __shared__ int s_unst, s_inc;
    // the total is the last inclusive prefix, so that target <= bincl[nb-1]
    if (b0 < nb && b0 + chunk >= nb) s_total = run;
    if (threadIdx.x == 1023) { s_unst = wu[15] + incl_u; s_inc = wi[15] + incl_i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        KmcCtl* kc = p.kctl;
        kc->total_rate = s_total;
        kc->n_unstable = s_unst;
        kc->n_included = s_inc;
        kc->nblocks = nb;
        kc->sel = -1;
        kc->ambiguous = 0;
        kc->ticket = 0;
        kc->go = 1;
        kc->outcome = 0;
    }
}

template <bool BATCH>
__global__ __launch_bounds__(1024) void k_rates_scan(const KmcParams p_arg, const KmcParams* __restrict__ PK, int nb) {
    rates_scan_body(param_sel<BATCH>(p_arg, PK), nb);
}

// ---------------------------------------------------------------------------
// Selection: block by a two-level search over the inclusive block prefixes
// (non-decreasing), mode by an in-block scan; a rounding-bound guard flags
// targets too close to call.
// ---------------------------------------------------------------------------
// The lanes 0..255 of the block work; every lane of a larger block takes
// part in the barriers.
__device__ __forceinline__ void select_body(const KmcParams& p, double u2) {
    const bool act = threadIdx.x < 256;
    const long nmodes = p.g.N * p.ph.M;
    KmcCtl* kc = p.kctl;
    if (!kc->go) return;
    const int nb = kc->nblocks;
    const double total = kc->total_rate;
    const double target = u2 * total;
    __shared__ int s_b, s_chunk;
    __shared__ double s_scan[kRateBlock];
    __shared__ double s_tot[256];
    __shared__ int s_first;
    if (threadIdx.x == 0) {
        s_b = nb - 1;   // no prefix reaches the target: the last block
        s_chunk = 0x7fffffff;
        s_first = kRateBlock;
    }
    __syncthreads();
    // first b with bincl[b] >= target: the first lane chunk whose last prefix
    // reaches it, then the first entry inside that chunk
    const int chunk = (nb + 255) / 256;
    {
        const int lo = threadIdx.x * chunk;
        const int last = min(nb, lo + chunk) - 1;
        if (act && lo < nb && p.bincl[last] >= target) atomicMin(&s_chunk, (int)threadIdx.x);
    }
    __syncthreads();
    if (act && s_chunk != 0x7fffffff) {
        for (int k = threadIdx.x; k < chunk; k += 256) {
            const int bb = s_chunk * chunk + k;
            if (bb < nb && p.bincl[bb] >= target) atomicMin(&s_b, bb);
        }
    }
    __syncthreads();
    const int b = s_b;
    const long base = (long)b * kRateBlock;
    const double prev = (b > 0) ? p.bincl[b - 1] : 0.0;
    // 4 consecutive modes per lane, local inclusive prefix
    double loc[4];
    bool inc[4];
    double run = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long i = base + threadIdx.x * 4 + k;
        bool incl = false;
        const double r = (act && i < nmodes) ? mode_rate(p.ph, p.Q[i], incl) : 0.0;
        run += incl ? r : 0.0;
        loc[k] = run;
        inc[k] = incl;
    }
    if (act) s_tot[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        double a = (act && (int)threadIdx.x >= off) ? s_tot[threadIdx.x - off] : 0.0;
        __syncthreads();
        if (act) s_tot[threadIdx.x] += a;
        __syncthreads();
    }
    const double excl = (act && threadIdx.x > 0) ? s_tot[threadIdx.x - 1] : 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int li = threadIdx.x * 4 + k;
        const long i = base + li;
        const double S = prev + (excl + loc[k]);
        if (!act) continue;
        s_scan[li] = S;
        if (i < nmodes && inc[k] && S >= target) atomicMin(&s_first, li);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double u = 1.1102230246251565e-16;   // 2^-53
    int sel = -1, amb = 1;
    if (s_first < kRateBlock) {
        const int li = s_first;
        const long i = base + li;
        const double S = s_scan[li];
        const double Sp = (li > 0) ? s_scan[li - 1] : prev;
        // reference prefix error <= i u S (left-to-right sum); device <= (log2 nb + 1024) u S;
        // target error <= nmodes u total
        const double mS = 2.0 * ((double)i + 2048.0) * u * S + 2.0 * (double)nmodes * u * total;
        const double mP = 2.0 * ((double)i + 2048.0) * u * Sp + 2.0 * (double)nmodes * u * total;
        sel = (int)i;
        amb = (S - target <= mS) || (target - Sp <= mP) ? 1 : 0;
    }
    kc->u2 = u2;
    kc->target = target;
    kc->sel = sel;
    kc->ambiguous = amb;
    kc->q_sel = (sel >= 0) ? p.Q[sel] : 0.0;
}

// The selection for a 1024-lane block that has just run rates_scan_body:
// each lane searches the inclusive prefixes of its own scan chunk (its own
// writes), and the chosen block's 1024 modes get one lane each, with a wave
// scan and a scan of the 16 wave totals.  The summation order differs from
// select_body's; the rounding-bound test holds for any order of the 1024
// terms.  It always names a mode when the total rate is positive:
//  * no in-block prefix reaches a target that the block's inclusive prefix
//    reaches (the two sums round differently): the block's last included
//    mode, flagged ambiguous;
//  * target 0 (u2 == 0) ahead of a block without included modes: the first
//    included mode after it, which is the reference's answer (every prefix of
//    an included mode is >= 0).
__device__ __forceinline__ void select_body_1024(const KmcParams& p, double u2) {
    const long nmodes = p.g.N * p.ph.M;
    KmcCtl* kc = p.kctl;
    if (!kc->go) return;
    const int nb = kc->nblocks;
    const double total = kc->total_rate;
    const double target = u2 * total;
    __shared__ int s_b1, s_first1, s_last1, s_fwd1;
    __shared__ double s_scan1[kRateBlock], s_q1[kRateBlock];
    __shared__ double s_wave[16];
    if (threadIdx.x == 0) {
        s_b1 = nb - 1;   // no prefix reaches the target: the last block
        s_first1 = kRateBlock;
        s_last1 = -1;
        s_fwd1 = -1;
    }
    __syncthreads();
    {
        // the first block of the lane's chunk (rates_scan_body's chunks) whose
        // inclusive prefix reaches the target; a group's prefixes load together
        constexpr int G = 8;
        const int chunk = (nb + 1023) / 1024;
        const int b0 = threadIdx.x * chunk;
        for (int g0 = 0; g0 < chunk; g0 += G) {
            double vb[G];
            bool ok[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const int b = b0 + g0 + k;
                ok[k] = (g0 + k < chunk) && (b < nb);
                vb[k] = p.bincl[ok[k] ? b : 0];
            }
            int hit = -1;
#pragma unroll
            for (int k = G - 1; k >= 0; --k)
                if (ok[k] && vb[k] >= target) hit = b0 + g0 + k;
            if (hit >= 0) {
                atomicMin(&s_b1, hit);
                break;
            }
        }
    }
    __syncthreads();
    const int b = s_b1;
    const long base = (long)b * kRateBlock;
    const double prev = (b > 0) ? p.bincl[b - 1] : 0.0;
    const long i = base + threadIdx.x;
    bool incl = false;
    const double qi = (i < nmodes) ? p.Q[i] : 0.0;
    s_q1[threadIdx.x] = qi;
    const double r = (i < nmodes) ? mode_rate(p.ph, qi, incl) : 0.0;
    double v = incl ? r : 0.0;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    if (lane == 63) s_wave[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int w = 0; w < 16; ++w) {
            const double t = s_wave[w];
            s_wave[w] = run;   // exclusive wave offsets
            run += t;
        }
    }
    __syncthreads();
    const double S = prev + (s_wave[wid] + v);
    s_scan1[threadIdx.x] = S;
    if (i < nmodes && incl) {
        if (S >= target) atomicMin(&s_first1, (int)threadIdx.x);
        atomicMax(&s_last1, (int)threadIdx.x);
    }
    __syncthreads();
    if (s_first1 == kRateBlock && s_last1 < 0) {
        // block b holds no included mode (only a target of 0 can pick such a
        // block): the first included mode of the blocks after it
        for (int bb = b + 1; bb < nb; ++bb) {
            const long j = (long)bb * kRateBlock + threadIdx.x;
            bool inc2 = false;
            if (j < nmodes) (void)mode_rate(p.ph, p.Q[j], inc2);
            if (inc2) atomicMin(&s_first1, (int)threadIdx.x);
            __syncthreads();
            const bool found = s_first1 < kRateBlock;
            if (found && threadIdx.x == 0) s_fwd1 = bb;
            __syncthreads();
            if (found) break;
        }
    }
    if (threadIdx.x != 0) return;
    const double u = 1.1102230246251565e-16;   // 2^-53
    int sel = -1, amb = 1;
    double qs = 0.0;
    if (s_fwd1 >= 0) {
        sel = (int)((long)s_fwd1 * kRateBlock + s_first1);
        amb = 0;
        qs = p.Q[sel];
    } else if (s_first1 == kRateBlock && s_last1 >= 0) {
        sel = (int)(base + s_last1);
        amb = 1;
        qs = s_q1[s_last1];
    } else if (s_first1 < kRateBlock) {
        const int li = s_first1;
        const long im = base + li;
        const double Sm = s_scan1[li];
        const double Sp = (li > 0) ? s_scan1[li - 1] : prev;
        const double mS = 2.0 * ((double)im + 2048.0) * u * Sm + 2.0 * (double)nmodes * u * total;
        const double mP = 2.0 * ((double)im + 2048.0) * u * Sp + 2.0 * (double)nmodes * u * total;
        sel = (int)im;
        amb = (Sm - target <= mS) || (target - Sp <= mP) ? 1 : 0;
        qs = s_q1[li];
    }
    kc->u2 = u2;
    kc->target = target;
    kc->sel = sel;
    kc->ambiguous = amb;
    kc->q_sel = qs;
}

template <bool BATCH>
__global__ __launch_bounds__(256) void k_select(const KmcParams p_arg, const KmcParams* __restrict__ PK, double u2,
                                                const EventArgs* __restrict__ EA) {
    if (BATCH) u2 = EA[blockIdx.y].u2;
    select_body(param_sel<BATCH>(p_arg, PK), u2);
}

// Rates of the kRateBlock modes of chunk b into dst, excluded modes as `excl`,
// by `lanes` lanes starting at lane index t.
__device__ __forceinline__ void chunk_rates(const KmcParams& p, int b, long nmodes, double excl, double* dst, int t,
                                            int lanes) {
    constexpr int MAXPER = kRateBlock / 64;
    double q[MAXPER];
#pragma unroll
    for (int k = 0; k < MAXPER; ++k) {
        const int j = t + k * lanes;
        const long i = (long)b * kRateBlock + j;
        q[k] = (j < kRateBlock && i < nmodes) ? p.Q[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < MAXPER; ++k) {
        const int j = t + k * lanes;
        if (j >= kRateBlock) break;
        const long i = (long)b * kRateBlock + j;
        bool incl = false;
        const double r = (i < nmodes) ? mode_rate(p.ph, q[k], incl) : 0.0;
        dst[j] = incl ? r : excl;
    }
}

// Exact left-to-right pass of the reference (kmc_solver.hpp:790-826, :1803-1845):
// the total in flat order, target = u2 * total, then the first included mode
// whose running sum reaches the target.  The additions are the only
// sequential part: wave 0 runs them in flat order over a chunk of kRateBlock
// rates in LDS (adding 0 for an excluded mode leaves the sum exact) while
// waves 1-3 evaluate the next chunk's rates into the other LDS buffer.  Pass 1
// keeps the running sum before every chunk in bsum (free once k_rates_scan
// has read it); pass 2 restarts at the first chunk whose end reaches the
// target.  With only_if_ambiguous it runs only when k_select flagged its answer.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_select_seq(const KmcParams p_arg, const KmcParams* __restrict__ PK,
                                                    int only_if_ambiguous) {
    const KmcParams& p = param_sel<BATCH>(p_arg, PK);
    KmcCtl* kc = p.kctl;
    if (!kc->go || (only_if_ambiguous && !kc->ambiguous)) return;
    __shared__ double rb[2][kRateBlock];
    __shared__ double s_tot;
    __shared__ int s_bs;
    __shared__ long s_sel;
    const long nmodes = p.g.N * p.ph.M;
    const int nb = (int)((nmodes + kRateBlock - 1) / kRateBlock);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int G = 16;   // rates read ahead of the adds
    if (wid > 0) chunk_rates(p, 0, nmodes, 0.0, rb[0], threadIdx.x - 64, 192);
    __syncthreads();
    double tot = 0.0;
    for (int b = 0; b < nb; ++b) {
        if (wid == 0) {
            if (lane == 0) p.bsum[b] = tot;
            const double* r = rb[b & 1];
            double cur[G], nxt[G];
#pragma unroll
            for (int k = 0; k < G; ++k) cur[k] = r[k];
            for (int g0 = 0; g0 < kRateBlock; g0 += G) {
                if (g0 + G < kRateBlock) {
#pragma unroll
                    for (int k = 0; k < G; ++k) nxt[k] = r[g0 + G + k];
                }
#pragma unroll
                for (int k = 0; k < G; ++k) tot += cur[k];
#pragma unroll
                for (int k = 0; k < G; ++k) cur[k] = nxt[k];
            }
        } else if (b + 1 < nb) {
            chunk_rates(p, b + 1, nmodes, 0.0, rb[(b + 1) & 1], threadIdx.x - 64, 192);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) s_tot = tot;
    __syncthreads();
    tot = s_tot;
    const double target = kc->u2 * tot;
    // first chunk whose end (the start of the next, or the total) reaches the target
    if (wid == 0) {
        int bs = nb;
        for (int b = lane; b < nb; b += 64) {
            const double e = (b + 1 < nb) ? p.bsum[b + 1] : tot;
            if (e >= target) { bs = b; break; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) bs = min(bs, __shfl_xor(bs, off, 64));
        // no chunk reaches a target above the total (u2 >= 1): scan everything,
        // ending at the last included mode
        if (lane == 0) { s_bs = (bs < nb) ? bs : 0; s_sel = -1; }
    }
    __syncthreads();
    const int bs = s_bs;
    double run = (bs < nb) ? p.bsum[bs] : 0.0;
    long last = -1;
    for (int b = bs; b < nb; ++b) {
        chunk_rates(p, b, nmodes, -1.0, rb[0], threadIdx.x, 256);
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int j = 0; j < kRateBlock; ++j) {
                const double r = rb[0][j];
                if (r < 0.0) continue;   // not included
                last = (long)b * kRateBlock + j;
                run += r;
                if (run >= target) { s_sel = last; break; }
            }
            if (s_sel < 0 && b == nb - 1) s_sel = last;   // the reference's default: the last included mode
        }
        __syncthreads();
        if (s_sel >= 0) break;
    }
    if (threadIdx.x == 0) {
        const long sel = s_sel;
        kc->target = target;
        kc->sel = (int)sel;
        kc->ambiguous = 0;
        kc->q_sel = (sel >= 0) ? p.Q[sel] : 0.0;
    }
}

// ---------------------------------------------------------------------------
// apply_flip_soa_2d/3d + prev_strain_dir + redraw upload (kmc_solver.hpp:423-462,
// :835-864; 3D :1424-1471, :1850-1875).  The neighbour loop runs in the
// reference's order on one lane: periodic wraps on small grids hit the same
// voxel several times and each hit accumulates.
// ---------------------------------------------------------------------------
template <int D, int C>
__device__ __forceinline__ void flip_body(const KmcParams& p, int flat, double time, int first_mode, int n_modes,
                                          int has_cat, int has_q0, int time_from_ctl, const RedrawArgs& ra, int ra_on,
                                          const double* stage_ptr = nullptr) {
    // flat < 0: the mode the last selection chose, if that attempt produced
    // an event; first_mode < 0: redraw only that mode (both redraw switches
    // off, kmc_solver.hpp:856-864)
    if (flat < 0) {
        if (!p.kctl->go) return;
        flat = p.kctl->sel;
    }
    if (flat < 0) return;
    if (time_from_ctl) time = p.kctl->time_new;
    if (first_mode < 0) first_mode = flat % p.ph.M;
    const Geom g = p.g;
    const KmcPhys ph = p.ph;
    const long nmodes = g.N * ph.M;
    const long v = flat / ph.M;
    const double* stage = stage_ptr ? stage_ptr : (ra_on ? ra.v : p.staging);
    double c[C];
#pragma unroll
    for (int k = 0; k < C; ++k) c[k] = p.cat_shared ? p.cat[k * ph.M + flat % ph.M] : p.cat[k * nmodes + flat];
#pragma unroll
    for (int k = 0; k < C; ++k) p.eps_p[k * g.N + v] += c[k];
#pragma unroll
    for (int k = 0; k < C; ++k) p.kctl->flip_c[k] = c[k];
    p.kctl->flip_voxel = (int)v;
    double sum_sq;
    if constexpr (D == 2) {
        const double e11 = c[0], e22 = c[1], e12 = c[2];
        sum_sq = (e12 * e12) + (e22 * e22 + e11 * e11 + (e11 - e22) * (e11 - e22)) / 6.0;
    } else {
        const double e11 = c[0], e22 = c[1], e33 = c[2], e12 = c[3], e13 = c[4], e23 = c[5];
        sum_sq = (e12 * e12 + e23 * e23 + e13 * e13) +
                 ((e22 - e33) * (e22 - e33) + (e33 - e11) * (e33 - e11) + (e11 - e22) * (e11 - e22)) / 6.0;
    }
    double gp = p.soft[v] + ph.jp * sum_sq;
    if (ph.cap > 0.0 && gp > ph.cap) gp = ph.cap;
    p.soft[v] = gp;
    p.soft[g.N + v] = ph.jt * sum_sq;
    p.last_t[v] = time;
    if (ph.nfrac > 0.0) {
        int x, y, z = 0;
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        if (D == 2) { y = (int)(v % ny); x = (int)(v / ny); }
        else { z = (int)(v % nz); const long t = v / nz; y = (int)(t % ny); x = (int)(t / ny); }
        constexpr int zr = (D == 3) ? 1 : 0;
        if (nx >= 3 && ny >= 3 && (D == 2 || nz >= 3)) {
            // every neighbour is a distinct voxel other than v: per dx plane,
            // load its neighbours together, then update them (the same
            // arithmetic per neighbour)
            constexpr int NP = (D == 2) ? 3 : 9;   // cells per dx plane
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                long nbs[NP];
                double g0[NP], g1[NP];
                const int xn = (x + dx + nx) % nx;
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const int dy = (D == 2) ? q - 1 : q / 3 - 1, dz = (D == 2) ? 0 : q % 3 - 1;
                    const int yn = (y + dy + ny) % ny;
                    const int zn = (D == 3) ? (z + dz + nz) % nz : 0;
                    nbs[q] = (D == 2) ? (long)xn * ny + yn : ((long)xn * ny + yn) * nz + zn;
                    const bool centre = (dx == 0 && dy == 0 && dz == 0);
                    g0[q] = p.soft[centre ? v : nbs[q]];
                    g1[q] = p.soft[g.N + (centre ? v : nbs[q])];
                }
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const int dy = (D == 2) ? q - 1 : q / 3 - 1, dz = (D == 2) ? 0 : q % 3 - 1;
                    if (dx == 0 && dy == 0 && dz == 0) continue;
                    double gn = g0[q] + ph.nfrac * ph.jp * sum_sq;
                    if (ph.cap > 0.0 && gn > ph.cap) gn = ph.cap;
                    p.soft[nbs[q]] = gn;
                    p.soft[g.N + nbs[q]] = g1[q] + ph.nfrac * ph.jt * sum_sq;
                }
            }
        } else {
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -zr; dz <= zr; ++dz) {
                        if (dx == 0 && dy == 0 && dz == 0) continue;
                        const int xn = (x + dx + nx) % nx, yn = (y + dy + ny) % ny;
                        const int zn = (D == 3) ? (z + dz + nz) % nz : 0;
                        const long nb = (D == 2) ? (long)xn * ny + yn : ((long)xn * ny + yn) * nz + zn;
                        double gn = p.soft[nb] + ph.nfrac * ph.jp * sum_sq;
                        if (ph.cap > 0.0 && gn > ph.cap) gn = ph.cap;
                        p.soft[nb] = gn;
                        p.soft[g.N + nb] += ph.nfrac * ph.jt * sum_sq;
                    }
        }
    }
#pragma unroll
    for (int k = 0; k < C; ++k) p.prev_dir[k * g.N + v] = c[k];
    for (int j = 0; j < n_modes; ++j) {
        const long mi = v * ph.M + first_mode + j;
        if (has_cat && !p.cat_shared)
            for (int k = 0; k < C; ++k) p.cat[k * nmodes + mi] = stage[j * C + k];
        if (has_q0) p.Q0[mi] = stage[n_modes * C + j];
    }
}

// flip_body spread over the lanes of a block (every lane calls it; it has
// barriers).  When every neighbour is a distinct voxel other than v (all axes
// of length >= 3), lane q < 3^D updates cell q of the neighbourhood, lanes
// 32..32+C-1 the voxel's eps_p and prev_dir components, and after a barrier
// (every lane has read the flipped mode's catalog row) the lanes write the
// redraw.  Each address is written by one lane with the serial version's
// arithmetic, so the memory state equals flip_body's.  Otherwise lane 0 runs
// flip_body.
template <int D, int C>
__device__ __forceinline__ void flip_body_block(const KmcParams& p, int flat, double time, int first_mode,
                                                int n_modes, int has_cat, int has_q0, int time_from_ctl,
                                                const RedrawArgs& ra, int ra_on, const double* stage_ptr = nullptr) {
    __shared__ int s_flat, s_first;
    __shared__ double s_time;
    const Geom g = p.g;
    const bool distinct = g.nx >= 3 && g.ny >= 3 && (D == 2 || g.nz >= 3) && blockDim.x >= 64;
    if (!distinct) {
        if (threadIdx.x == 0)
            flip_body<D, C>(p, flat, time, first_mode, n_modes, has_cat, has_q0, time_from_ctl, ra, ra_on, stage_ptr);
        return;
    }
    if (threadIdx.x == 0) {
        if (flat < 0) flat = p.kctl->go ? p.kctl->sel : -1;
        s_flat = flat;
        s_time = (flat >= 0 && time_from_ctl) ? p.kctl->time_new : time;
        s_first = (flat >= 0 && first_mode < 0) ? flat % p.ph.M : first_mode;
    }
    __syncthreads();
    flat = s_flat;
    if (flat < 0) return;   // uniform: no event
    time = s_time;
    first_mode = s_first;
    const KmcPhys ph = p.ph;
    const long nmodes = g.N * ph.M;
    const long v = flat / ph.M;
    const double* stage = stage_ptr ? stage_ptr : (ra_on ? ra.v : p.staging);
    const int t = threadIdx.x;
    constexpr int NCELL = (D == 2) ? 9 : 27;
    constexpr int CENTRE = NCELL / 2;
    // loads: the flipped mode's catalog row (every lane), the lane's cell,
    // the lane's eps_p component
    double c[C];
#pragma unroll
    for (int k = 0; k < C; ++k) c[k] = p.cat_shared ? p.cat[k * ph.M + flat % ph.M] : p.cat[k * nmodes + flat];
    long nb = v;
    double g0 = 0.0, g1 = 0.0;
    if (t < NCELL) {
        int x, y, z = 0;
        if (D == 2) { y = (int)(v % g.ny); x = (int)(v / g.ny); }
        else { z = (int)(v % g.nz); const long tt = v / g.nz; y = (int)(tt % g.ny); x = (int)(tt / g.ny); }
        const int dx = (D == 2) ? t / 3 - 1 : t / 9 - 1;
        const int dy = (D == 2) ? t % 3 - 1 : (t / 3) % 3 - 1;
        const int dz = (D == 2) ? 0 : t % 3 - 1;
        const int xn = (x + dx + g.nx) % g.nx, yn = (y + dy + g.ny) % g.ny;
        const int zn = (D == 3) ? (z + dz + g.nz) % g.nz : 0;
        nb = (D == 2) ? (long)xn * g.ny + yn : ((long)xn * g.ny + yn) * g.nz + zn;
        g0 = p.soft[nb];
        g1 = p.soft[g.N + nb];
    }
    const int kc = t - 32;
    double ep = 0.0;
    if (kc >= 0 && kc < C) ep = p.eps_p[kc * g.N + v];
    double sum_sq;
    if constexpr (D == 2) {
        const double e11 = c[0], e22 = c[1], e12 = c[2];
        sum_sq = (e12 * e12) + (e22 * e22 + e11 * e11 + (e11 - e22) * (e11 - e22)) / 6.0;
    } else {
        const double e11 = c[0], e22 = c[1], e33 = c[2], e12 = c[3], e13 = c[4], e23 = c[5];
        sum_sq = (e12 * e12 + e23 * e23 + e13 * e13) +
                 ((e22 - e33) * (e22 - e33) + (e33 - e11) * (e33 - e11) + (e11 - e22) * (e11 - e22)) / 6.0;
    }
    if (t == CENTRE) {
        double gp = g0 + ph.jp * sum_sq;
        if (ph.cap > 0.0 && gp > ph.cap) gp = ph.cap;
        p.soft[v] = gp;
        p.soft[g.N + v] = ph.jt * sum_sq;
        p.last_t[v] = time;
    } else if (t < NCELL && ph.nfrac > 0.0) {
        double gn = g0 + ph.nfrac * ph.jp * sum_sq;
        if (ph.cap > 0.0 && gn > ph.cap) gn = ph.cap;
        p.soft[nb] = gn;
        p.soft[g.N + nb] = g1 + ph.nfrac * ph.jt * sum_sq;
    }
    if (kc >= 0 && kc < C) {
#pragma unroll
        for (int k = 0; k < C; ++k)
            if (k == kc) {
                p.eps_p[k * g.N + v] = ep + c[k];
                p.prev_dir[k * g.N + v] = c[k];
                p.kctl->flip_c[k] = c[k];
            }
        if (kc == 0) p.kctl->flip_voxel = (int)v;
    }
    __syncthreads();   // the catalog row is read before the redraw may overwrite it
    for (int j = t; j < n_modes; j += blockDim.x) {
        const long mi = v * ph.M + first_mode + j;
        if (has_cat && !p.cat_shared)
            for (int k = 0; k < C; ++k) p.cat[k * nmodes + mi] = stage[j * C + k];
        if (has_q0) p.Q0[mi] = stage[n_modes * C + j];
    }
}

// A cascade round's flips in their order (run_cascade, kmc_solver.hpp:488-523,
// :1500-1533): one block applies them one after another, each flip's
// neighbourhood updates complete before the next flip reads it.  Flip i's
// redraw values sit at stage + i * n_modes * (C + 1) (catalog then Q0); every
// flip redraws all M modes (all_modes) or only its own mode.
template <int D, int C>
__global__ void k_flip_list(const KmcParams p, const int* __restrict__ flats, const double* __restrict__ stage, int n,
                            double time, int all_modes, int has_cat, int has_q0) {
    if (blockIdx.x != 0) return;
    const int n_modes = all_modes ? p.ph.M : 1;
    const RedrawArgs ra{};
    for (int i = 0; i < n; ++i) {
        flip_body_block<D, C>(p, flats[i], time, all_modes ? 0 : -1, n_modes, has_cat, has_q0, 0, ra, 0,
                              stage + (size_t)i * n_modes * (C + 1));
        __syncthreads();
    }
}

template <int D, int C, bool BATCH>
__global__ void k_flip(const KmcParams p_arg, const KmcParams* __restrict__ PK, int flat, double time, int first_mode,
                       int n_modes, int has_cat, int has_q0, int time_from_ctl, const RedrawArgs ra, int ra_on) {
    if (blockIdx.x != 0) return;
    flip_body_block<D, C>(param_sel<BATCH>(p_arg, PK), flat, time, first_mode, n_modes, has_cat, has_q0, time_from_ctl,
                          ra, ra_on);
}

// ---------------------------------------------------------------------------
// find_unstable_indices (kmc_solver.hpp:412-421): ascending flat indices with
// Q < threshold, compacted with the block offsets of the rate pass.
// ---------------------------------------------------------------------------
template <bool BATCH>
__global__ __launch_bounds__(256) void k_unstable(const KmcParams p_arg, const KmcParams* __restrict__ PK) {
    const KmcParams& p = param_sel<BATCH>(p_arg, PK);
    const long nmodes = p.g.N * p.ph.M;
    const long base = (long)blockIdx.x * kRateBlock;
    __shared__ int s_cnt[256];
    int flags[4], cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long i = base + threadIdx.x * 4 + k;
        flags[k] = (i < nmodes && p.Q[i] < p.ph.thr) ? 1 : 0;
        cnt += flags[k];
    }
    s_cnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int a = ((int)threadIdx.x >= off) ? s_cnt[threadIdx.x - off] : 0;
        __syncthreads();
        s_cnt[threadIdx.x] += a;
        __syncthreads();
    }
    int pos = p.bunst_off[blockIdx.x] + ((threadIdx.x > 0) ? s_cnt[threadIdx.x - 1] : 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!flags[k]) continue;
        const long i = base + threadIdx.x * 4 + k;
        p.unst_idx[pos] = (int)i;
        p.unst_q[pos] = p.Q[i];
        pos++;
    }
}

// ---------------------------------------------------------------------------
// One attempt of the reference's inner loop decided on the device
// (kmc_solver.hpp:763-815, :942-952; 3D :1776-1826, :1967-1977), after the
// rate pass:
//   instability_mode "cascade" with unstable modes  -> MGK_EVENT_UNSTABLE
//   total rate not > 0                              -> MGK_EVENT_NO_RATE
//   t_wait = -log(u1) / R not < remaining           -> MGK_EVENT_REJECTED
//   else                                            -> MGK_EVENT_DONE:
//     time_new = time + t_wait, eps_macro(drv) += strain_rate * t_wait, and
//     the LS control block gets the imposed strain (symmetric part, norm of
//     the antisymmetric part) and the warm-start shift.
// Every operation is the host's own, in its order (no contraction in this
// file), so time_new and the imposed strain equal the host's bit for bit.
// ---------------------------------------------------------------------------
template <int D, int C>
__device__ __forceinline__ void event_prep_body(const KmcParams& p, Ctl* ctl, const EventArgs& ea) {
    KmcCtl* kc = p.kctl;
    if (ea.init_ctl) {
        // the host's control-block reset for a solve: zero, then max_iter, tol,
        // the field mean and the warm-start mode
        double* w = reinterpret_cast<double*>(ctl);
        for (int k = 0; k < (int)(sizeof(Ctl) / sizeof(double)); ++k) w[k] = 0.0;
        ctl->max_iter = ea.max_iter;
        ctl->cycle_len = ea.cycle_len;
        ctl->cycle_len2 = ea.cycle_len2;
        ctl->tol = ea.tol;
        ctl->keep_shift = ea.keep_shift;
        for (int c = 0; c < 6; ++c) ctl->mean_eps[c] = ea.mean_eps[c];
    }
    const double R = kc->total_rate;
    int outcome;
    double tw = 0.0;
    if (!ea.active) outcome = 4;                                      // MGK_EVENT_SKIPPED
    else if (ea.cascade_check && kc->n_unstable > 0) outcome = 1;     // MGK_EVENT_UNSTABLE
    else if (!(R > 0.0)) outcome = 2;                                 // MGK_EVENT_NO_RATE
    else {
        tw = ea.neg_log_u1 / R;
        outcome = (tw < ea.remaining) ? 0 : 3;                        // MGK_EVENT_DONE / _REJECTED
    }
    kc->t_wait = tw;
    kc->outcome = outcome;
    kc->go = (outcome == 0) ? 1 : 0;
    if (outcome != 0) {
        ctl->done = 1;
        ctl->skip_final = 1;
        return;
    }
    kc->time_new = ea.time + tw;
    // eps_macro += strain_unit * (strain_rate * t_wait), entry by entry
    const double x = ea.strain_rate * tw;
    double a[D * D];
#pragma unroll
    for (int k = 0; k < D * D; ++k) a[k] = ea.a[k] + ((k == ea.drv) ? 1.0 : 0.0) * x;
    double sym[C], skew[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = i; j < D; ++j) {
            const int c = comp_of(D, i, j);
            if (i == j) sym[c] = a[i + D * i];
            else {
                sym[c] = 0.5 * (a[i + D * j] + a[j + D * i]);
                const int w = (D == 2) ? 0 : (c - 3);
                skew[w] = 0.5 * (a[i + D * j] - a[j + D * i]);
            }
        }
    double norm_extra = 0.0;
    constexpr int NS = (D == 2) ? 1 : 3;
#pragma unroll
    for (int w = 0; w < NS; ++w) norm_extra += 2.0 * ea.n_vox * skew[w] * skew[w];
    ctl->norm_extra = norm_extra;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        ctl->eps_bar[c] = sym[c];
        if (ea.keep_shift) ctl->shift[c] = sym[c] - ctl->mean_eps[c];
    }
}

// ---------------------------------------------------------------------------
// One block per trajectory for everything between the rate pass and the LS
// batch of a device-decided event: the block-sum scan, the event decision,
// the selection and (FLIP) the flip with its redraw, in one launch (the
// selection as select_body_1024).  FP64 contexts leave the flip out:
// the exact selection pass runs between.
// ---------------------------------------------------------------------------
template <int D, int C, bool BATCH, bool FLIP>
__global__ __launch_bounds__(1024) void k_event_chain(const KmcParams p_arg, const KmcParams* __restrict__ PK,
                                                      Ctl* ctl_arg, const LsParams* __restrict__ P,
                                                      const EventArgs ea_arg, const EventArgs* __restrict__ EA, int nb,
                                                      int first_mode, int n_modes, int has_cat, int has_q0,
                                                      const RedrawArgs ra, int ra_on) {
    const KmcParams& p = param_sel<BATCH>(p_arg, PK);
    const EventArgs& ea = param_sel<BATCH>(ea_arg, EA);
    rates_scan_body(p, nb);
    __syncthreads();
    if (threadIdx.x == 0) event_prep_body<D, C>(p, BATCH ? P[blockIdx.y].ctl : ctl_arg, ea);
    __syncthreads();
    select_body_1024(p, ea.u2);
    if constexpr (FLIP) {
        __syncthreads();
        if (threadIdx.x == 0 && p.kctl->go && p.kctl->sel < 0) {
            // unreachable with a positive total rate; never report an event
            // without a mode: no solve, no final pass, outcome "no rate"
            Ctl* ctl = BATCH ? P[blockIdx.y].ctl : ctl_arg;
            p.kctl->go = 0;
            p.kctl->outcome = 2;
            ctl->done = 1;
            ctl->skip_final = 1;
        }
        __syncthreads();
        flip_body_block<D, C>(p, -1, ea.time, first_mode, n_modes, has_cat, has_q0, 1, ra, ra_on);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int kmc_nblocks(long nmodes) { return (int)((nmodes + kRateBlock - 1) / kRateBlock); }

template <bool BATCH> static void enqueue_rates_t(const KmcLaunch& l, double time, const double* times, hipStream_t s,
                                                  bool with_scan) {
    const int nb = kmc_nblocks(l.g.N * l.M);
    const bool staged = l.M >= kStageMinM;
    if (l.g.dims == 2) {
        if (staged)
            hipLaunchKernelGGL((k_rates_staged<2, 3, BATCH>), dim3(nb, l.B), dim3(kRateThreads), 0, s, l.h, l.P, time, times);
        else
            hipLaunchKernelGGL((k_rates<2, 3, BATCH>), dim3(nb, l.B), dim3(kRateThreads), 0, s, l.h, l.P, time, times);
    } else {
        if (staged)
            hipLaunchKernelGGL((k_rates_staged<3, 6, BATCH>), dim3(nb, l.B), dim3(kRateThreads), 0, s, l.h, l.P, time, times);
        else
            hipLaunchKernelGGL((k_rates<3, 6, BATCH>), dim3(nb, l.B), dim3(kRateThreads), 0, s, l.h, l.P, time, times);
    }
    if (with_scan) hipLaunchKernelGGL(k_rates_scan<BATCH>, dim3(1, l.B), dim3(1024), 0, s, l.h, l.P, nb);
}

void kmc_enqueue_rates(const KmcLaunch& l, double time, const double* times, hipStream_t s, bool with_scan) {
    if (l.batch) {
        if (!times) throw ArgError("a batch needs per-trajectory times");
        enqueue_rates_t<true>(l, time, times, s, with_scan);
    } else {
        enqueue_rates_t<false>(l, time, times, s, with_scan);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

void kmc_enqueue_select(const KmcLaunch& l, double u2, const EventArgs* ea, hipStream_t s) {
    if (l.batch) {
        if (!ea) throw ArgError("a batch needs per-trajectory event arguments");
        hipLaunchKernelGGL(k_select<true>, dim3(1, l.B), dim3(256), 0, s, l.h, l.P, u2, ea);
    } else {
        hipLaunchKernelGGL(k_select<false>, dim3(1, 1), dim3(256), 0, s, l.h, l.P, u2, ea);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

void kmc_enqueue_select_fallback(const KmcLaunch& l, int only_if_ambiguous, hipStream_t s) {
    if (l.batch) hipLaunchKernelGGL(k_select_seq<true>, dim3(1, l.B), dim3(256), 0, s, l.h, l.P, only_if_ambiguous);
    else hipLaunchKernelGGL(k_select_seq<false>, dim3(1, 1), dim3(256), 0, s, l.h, l.P, only_if_ambiguous);
    MGK_HIP_CHECK(hipGetLastError());
}

template <int D, int C, bool BATCH>
static void enqueue_flip_t(const KmcLaunch& l, int flat, double time, int first_mode, int n_modes, int has_cat,
                           int has_q0, hipStream_t s, int time_from_ctl, const RedrawArgs& ra, int ra_on) {
    hipLaunchKernelGGL((k_flip<D, C, BATCH>), dim3(1, l.B), dim3(64), 0, s, l.h, l.P, flat, time, first_mode, n_modes,
                       has_cat, has_q0, time_from_ctl, ra, ra_on);
}

bool kmc_redraw_fits_args(int n_modes, int C) { return n_modes * (C + 1) <= kRedrawArgs; }

void kmc_enqueue_flip_list(const KmcLaunch& l, const int* flats, const double* stage, int n, double time,
                           int all_modes, int has_cat, int has_q0, hipStream_t s) {
    if (l.batch) throw ArgError("flip lists run on a single trajectory");
    if (l.g.dims == 2)
        hipLaunchKernelGGL((k_flip_list<2, 3>), dim3(1), dim3(64), 0, s, l.h, flats, stage, n, time, all_modes, has_cat,
                           has_q0);
    else
        hipLaunchKernelGGL((k_flip_list<3, 6>), dim3(1), dim3(64), 0, s, l.h, flats, stage, n, time, all_modes, has_cat,
                           has_q0);
    MGK_HIP_CHECK(hipGetLastError());
}

void kmc_enqueue_flip(const KmcLaunch& l, int flat, double time, int first_mode, int n_modes, int has_cat,
                      int has_q0, hipStream_t s, int time_from_ctl, const double* host_redraw) {
    if (l.batch && (flat >= 0 || !time_from_ctl)) throw ArgError("a batch flips its device-selected modes");
    RedrawArgs ra{};
    int ra_on = 0;
    if (host_redraw) {
        if (l.batch || !kmc_redraw_fits_args(n_modes, l.g.C)) throw ArgError("redraw values too many for kernel arguments");
        for (int k = 0; k < n_modes * (l.g.C + 1); ++k) ra.v[k] = host_redraw[k];
        ra_on = 1;
    }
    if (l.g.dims == 2) {
        if (l.batch) enqueue_flip_t<2, 3, true>(l, flat, time, first_mode, n_modes, has_cat, has_q0, s, time_from_ctl, ra, 0);
        else enqueue_flip_t<2, 3, false>(l, flat, time, first_mode, n_modes, has_cat, has_q0, s, time_from_ctl, ra, ra_on);
    } else {
        if (l.batch) enqueue_flip_t<3, 6, true>(l, flat, time, first_mode, n_modes, has_cat, has_q0, s, time_from_ctl, ra, 0);
        else enqueue_flip_t<3, 6, false>(l, flat, time, first_mode, n_modes, has_cat, has_q0, s, time_from_ctl, ra, ra_on);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

template <int D, int C, bool BATCH>
static void enqueue_chain_t(const KmcLaunch& l, Ctl* ctl, const LsParams* P, const EventArgs& ea_one,
                            const EventArgs* ea, bool flip, int first_mode, int n_modes, int has_cat, int has_q0,
                            const RedrawArgs& ra, int ra_on, hipStream_t s) {
    const int nb = kmc_nblocks(l.g.N * l.M);
    if (flip)
        hipLaunchKernelGGL((k_event_chain<D, C, BATCH, true>), dim3(1, l.B), dim3(1024), 0, s, l.h, l.P, ctl, P, ea_one,
                           ea, nb, first_mode, n_modes, has_cat, has_q0, ra, ra_on);
    else
        hipLaunchKernelGGL((k_event_chain<D, C, BATCH, false>), dim3(1, l.B), dim3(1024), 0, s, l.h, l.P, ctl, P,
                           ea_one, ea, nb, first_mode, n_modes, has_cat, has_q0, ra, ra_on);
}

void kmc_enqueue_event_chain(const KmcLaunch& l, Ctl* ctl, const LsParams* P, const EventArgs& ea_one,
                             const EventArgs* ea, bool flip, int first_mode, int n_modes, int has_cat, int has_q0,
                             const double* host_redraw, hipStream_t s) {
    if (l.batch ? (!P || !ea) : (!ctl || l.B != 1)) throw ArgError("event decision: missing control blocks");
    RedrawArgs ra{};
    int ra_on = 0;
    if (host_redraw) {
        if (l.batch || !kmc_redraw_fits_args(n_modes, l.g.C)) throw ArgError("redraw values too many for kernel arguments");
        for (int k = 0; k < n_modes * (l.g.C + 1); ++k) ra.v[k] = host_redraw[k];
        ra_on = 1;
    }
    if (l.g.dims == 2) {
        if (l.batch) enqueue_chain_t<2, 3, true>(l, ctl, P, ea_one, ea, flip, first_mode, n_modes, has_cat, has_q0, ra, 0, s);
        else enqueue_chain_t<2, 3, false>(l, ctl, P, ea_one, ea, flip, first_mode, n_modes, has_cat, has_q0, ra, ra_on, s);
    } else {
        if (l.batch) enqueue_chain_t<3, 6, true>(l, ctl, P, ea_one, ea, flip, first_mode, n_modes, has_cat, has_q0, ra, 0, s);
        else enqueue_chain_t<3, 6, false>(l, ctl, P, ea_one, ea, flip, first_mode, n_modes, has_cat, has_q0, ra, ra_on, s);
    }
    MGK_HIP_CHECK(hipGetLastError());
}


// ---------------------------------------------------------------------------
// Fast patching (kmc_solver.hpp:525-547, :866-913, :941-949; 3D :1535-1573,
// :1877-1939, :1967-1972).  One lane per voxel applies, in the reference's
// order: the predictor sig += sigma_unit * (strain_rate * t_wait), the window
// cells whose periodic image is this voxel (dx, dy, dz ascending, as the
// reference's loops meet them), then the missing-mean shift.  Basis
// coefficients: 2D (c_xx, c_xy); 3D (c_xx, -c_zz, c_xy, c_xz, c_yz).
// ---------------------------------------------------------------------------
// Offsets d in [-R, R] with (x0 + d) mod n == x, ascending: the first and
// their count (d = r + k n with r = (x - x0) mod n).
__device__ __forceinline__ void window_hits(int x0, int x, int n, int R, int& d0, int& cnt) {
    int r = (x - x0) % n;
    if (r < 0) r += n;
    d0 = r - ((R + r) / n) * n;   // smallest k with r + k n >= -R
    cnt = (d0 <= R) ? (R - d0) / n + 1 : 0;
}

template <int D, int C>
__global__ __launch_bounds__(256) void k_patch(const KmcParams p, const PatchParams pp, int voxel, double c0, double c1,
                                               double c2, double c3, double c4, double c5, double pred, int from_ctl,
                                               double strain_rate) {
    const Geom g = p.g;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.N) return;
    double cs[6] = {c0, c1, c2, c3, c4, c5};
    if (from_ctl) {
        const KmcCtl* kc = p.kctl;
        if (!kc->go) return;
        voxel = kc->flip_voxel;
#pragma unroll
        for (int k = 0; k < C; ++k) cs[k] = kc->flip_c[k];
        pred = (kc->t_wait > 0.0) ? strain_rate * kc->t_wait : 0.0;
    }
    double a[5];
    if constexpr (D == 2) { a[0] = cs[0]; a[1] = cs[2]; }
    else { a[0] = cs[0]; a[1] = -cs[2]; a[2] = cs[3]; a[3] = cs[4]; a[4] = cs[5]; }
    constexpr int NB = (D == 2) ? 2 : 5;
    double s[C];
#pragma unroll
    for (int k = 0; k < C; ++k) s[k] = p.sig[k * g.N + v];
    if (pred != 0.0) {
#pragma unroll
        for (int k = 0; k < C; ++k) s[k] += pp.sunit[k] * pred;
    }
    if (voxel >= 0) {
        int x, y, z = 0, x0, y0, z0 = 0;
        if (D == 2) { y = (int)(v % g.ny); x = (int)(v / g.ny); y0 = voxel % g.ny; x0 = voxel / g.ny; }
        else {
            z = (int)(v % g.nz); const long t = v / g.nz; y = (int)(t % g.ny); x = (int)(t / g.ny);
            z0 = voxel % g.nz; const int t0 = voxel / g.nz; y0 = t0 % g.ny; x0 = t0 / g.ny;
        }
        int dx0, nxh, dy0, nyh, dz0 = 0, nzh = 1;
        window_hits(x0, x, g.nx, pp.R, dx0, nxh);
        window_hits(y0, y, g.ny, pp.R, dy0, nyh);
        if (D == 3) window_hits(z0, z, g.nz, pp.zr, dz0, nzh);
        for (int ix = 0; ix < nxh; ++ix)
            for (int iy = 0; iy < nyh; ++iy)
                for (int iz = 0; iz < nzh; ++iz) {
                    const int dx = dx0 + ix * g.nx, dy = dy0 + iy * g.ny, dz = dz0 + iz * g.nz;
                    const int w = ((dx + pp.R) * pp.side + (dy + pp.R)) * pp.side_z + (dz + pp.zr);
#pragma unroll
                    for (int k = 0; k < C; ++k) {
                        double pv = a[0] * pp.K[(0 * C + k) * pp.W + w];
#pragma unroll
                        for (int b = 1; b < NB; ++b) pv += a[b] * pp.K[(b * C + k) * pp.W + w];
                        s[k] += pv;
                    }
                }
#pragma unroll
        for (int k = 0; k < C; ++k) {
            double m = a[0] * pp.Km[0 * C + k];
#pragma unroll
            for (int b = 1; b < NB; ++b) m += a[b] * pp.Km[b * C + k];
            s[k] += m;
        }
    }
#pragma unroll
    for (int k = 0; k < C; ++k) const_cast<double*>(p.sig)[k * g.N + v] = s[k];   // the one pass that writes sig
}

void kmc_enqueue_patch(const KmcLaunch& l, const PatchParams& pp, int voxel, const double* c, double pred,
                       int from_ctl, double strain_rate, hipStream_t s) {
    if (l.batch) throw ArgError("fast patching runs on single trajectories");
    double cs[6] = {0, 0, 0, 0, 0, 0};
    if (c)
        for (int k = 0; k < l.g.C; ++k) cs[k] = c[k];
    const int nb = (int)((l.g.N + 255) / 256);
    if (l.g.dims == 2)
        hipLaunchKernelGGL((k_patch<2, 3>), dim3(nb), dim3(256), 0, s, l.h, pp, voxel, cs[0], cs[1], cs[2], cs[3],
                           cs[4], cs[5], pred, from_ctl, strain_rate);
    else
        hipLaunchKernelGGL((k_patch<3, 6>), dim3(nb), dim3(256), 0, s, l.h, pp, voxel, cs[0], cs[1], cs[2], cs[3],
                           cs[4], cs[5], pred, from_ctl, strain_rate);
    MGK_HIP_CHECK(hipGetLastError());
}

void kmc_enqueue_unstable(const KmcLaunch& l, hipStream_t s) {
    const int nb = kmc_nblocks(l.g.N * l.M);
    if (l.batch) hipLaunchKernelGGL(k_unstable<true>, dim3(nb, l.B), dim3(256), 0, s, l.h, l.P);
    else hipLaunchKernelGGL(k_unstable<false>, dim3(nb, 1), dim3(256), 0, s, l.h, l.P);
    MGK_HIP_CHECK(hipGetLastError());
}

}  // namespace mgk
