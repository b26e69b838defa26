// ls_kernels.hip -- FFT Lippmann-Schwinger basic scheme on gfx950.
//
// One iteration of spectral_solver_2d/3d (linear_elastic_solver.hpp:203-297,
// :494-581) is a pipeline of LDS-resident FFT passes over the half spectrum
// (real-to-complex along the contiguous last axis L):
//
//   A  fwd_last   : tau = C(x):(eps - eps_p) - C0:eps per voxel, r2c along L
//                   (two real lines packed into one complex line)
//   B  mid   fwd  : c2c along M (3D only)
//   C  first_gamma: c2c along F, eps_hat = -Gamma0_eff(xi):tau_hat / N on the
//                   fly (no stored Green tensor), inverse c2c along F
//   D  mid   inv  : c2c along M (3D only)
//   E  inv_last   : c2r along L, eps_new = eps_bar + eps_tilde, the
//                   relative-increment norms, and a last-block reduction that
//                   sets the convergence flag in the control block
//
// Every kernel returns at once when the control block says the solve is done,
// so the host enqueues iterations in batches without a sync per iteration.
//
// Spectrum layout: component-interleaved half spectrum,
//   spec[((f*nm + m)*nlh + l)*C + c]   (complex),
// so every pass moves contiguous rows of (bin, component) pairs.
//
// The same passes run in two arithmetic modes:
//   FULL  (FP64 spectra): the reference's iteration on eps.
//   CG    (FP32 spectra): one conjugate-gradient step (Chronopoulos-Gear
//         form, one fused reduction) on A d = r with A = I + Gamma0*(C - C0),
//         SPD on compatible fields in the C0-energy inner product: pass A
//         updates the FP32 vectors (p, s, x, r) and forms tau = (C - C0):r,
//         pass E forms w = A r and the products (r,r)_C0, (w,r)_C0, ||r||_W.
//         This is the inner solver of the FP32 iterative-refinement path.
//
// Half-spectrum equivalence.  The reference transforms real fields with c2c
// and keeps the real part of the inverse (linear_elastic_solver.hpp:134-143).
// Re(IFFT(X)) equals the c2r transform of the Hermitian part of X.  For
// X = -Gamma0(xi_k) tau_hat(k) that part is -Gamma0_eff(k) tau_hat(k) with
// Gamma0_eff(k) = (Gamma0(xi_k) + Gamma0(xi_{-k}))/2, which differs from
// Gamma0(xi_k) only on Nyquist planes of even axes: fftfreq gives the Nyquist
// index the frequency -n/2 for both k and -k (linear_elastic_solver.hpp:22-33),
// so there xi_{-k} is xi_k with its Nyquist components' signs flipped.  The
// C pass evaluates both and averages them there.
//
// Mean subtraction (linear_elastic_solver.hpp:263-271): mean(eps_tilde) is the
// inverse transform's zero bin, which is exactly 0 because Gamma0(0) = 0
// (:86-94); the reference subtracts its rounding residue (~1e-17 |eps|).  The
// device does not (an O(ulp) difference).
#include "common.hpp"
#include "fft_lds.hpp"
#include "ls_kernels.hpp"
#include <algorithm>

namespace mgk {

extern __shared__ __align__(16) unsigned char smem_raw[];

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Deterministic block sum of K doubles per thread; result valid in thread 0.
template <int K> __device__ __forceinline__ void block_sum(double (&v)[K], double* scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) scratch[wid * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += scratch[w * K + k];
            v[k] = s;
        }
    }
    __syncthreads();
}

// Fixed-order reduction of part[b*K + k] over blocks (one block, plain loads:
// the partials were written by an earlier kernel on the same stream).
template <int K> __device__ __forceinline__ void reduce_partials(const double* part, int nblocks, double (&out)[K],
                                                                 double* scratch) {
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += part[(long)b * K + k];
    block_sum<K>(v, scratch);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = v[k];
}

template <typename SR> __device__ __forceinline__ void stage_twiddles(cplx<SR>* dst, const cplx<SR>* src, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Padded LDS line length (complex elements): keeps lines of neighbouring
// lanes on different bank groups.
template <int N> constexpr int padded() { return N + ((N % 16 == 0) ? 1 : 0); }

// ---------------------------------------------------------------------------
// Green operator on the fly (linear_elastic_solver.hpp:70-119, :344-396)
// With unit direction q and tau symmetric:
//   (Gamma0:tau)_kh = (q_h v_k + q_k v_h)/(2 mu0) - B q_k q_h (q.v),  v = tau q,
//   B = (lam0 + mu0)/(mu0 (lam0 + 2 mu0)).
// Gamma0 is homogeneous of degree 0 in xi, so q = (k_a / n_a)/|.| is used.
// ---------------------------------------------------------------------------
template <int D, typename R>
__device__ __forceinline__ void gamma_apply(const cplx<R>* t, cplx<R>* o, const R* q, R a, R B) {
    if constexpr (D == 2) {
        // t: xx, yy, xy
        cplx<R> vx = {t[0].x * q[0] + t[2].x * q[1], t[0].y * q[0] + t[2].y * q[1]};
        cplx<R> vy = {t[2].x * q[0] + t[1].x * q[1], t[2].y * q[0] + t[1].y * q[1]};
        cplx<R> s = {q[0] * vx.x + q[1] * vy.x, q[0] * vx.y + q[1] * vy.y};
        o[0] = {(R)2 * q[0] * vx.x * a - B * q[0] * q[0] * s.x, (R)2 * q[0] * vx.y * a - B * q[0] * q[0] * s.y};
        o[1] = {(R)2 * q[1] * vy.x * a - B * q[1] * q[1] * s.x, (R)2 * q[1] * vy.y * a - B * q[1] * q[1] * s.y};
        o[2] = {(q[1] * vx.x + q[0] * vy.x) * a - B * q[0] * q[1] * s.x,
                (q[1] * vx.y + q[0] * vy.y) * a - B * q[0] * q[1] * s.y};
    } else {
        // t: xx, yy, zz, xy, xz, yz
        cplx<R> vx = {t[0].x * q[0] + t[3].x * q[1] + t[4].x * q[2], t[0].y * q[0] + t[3].y * q[1] + t[4].y * q[2]};
        cplx<R> vy = {t[3].x * q[0] + t[1].x * q[1] + t[5].x * q[2], t[3].y * q[0] + t[1].y * q[1] + t[5].y * q[2]};
        cplx<R> vz = {t[4].x * q[0] + t[5].x * q[1] + t[2].x * q[2], t[4].y * q[0] + t[5].y * q[1] + t[2].y * q[2]};
        cplx<R> s = {q[0] * vx.x + q[1] * vy.x + q[2] * vz.x, q[0] * vx.y + q[1] * vy.y + q[2] * vz.y};
        const cplx<R> v[3] = {vx, vy, vz};
        const int ii[6] = {0, 1, 2, 0, 0, 1};
        const int jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const int i = ii[c], j = jj[c];
            o[c] = {(q[j] * v[i].x + q[i] * v[j].x) * a - B * q[i] * q[j] * s.x,
                    (q[j] * v[i].y + q[i] * v[j].y) * a - B * q[i] * q[j] * s.y};
        }
    }
}

// ---------------------------------------------------------------------------
// Pass A: tau + r2c along L.  Lines l = f*nm + m, LPW lines per block.
// ---------------------------------------------------------------------------
template <typename SR, int NL, int C, int MODE>
__global__ __launch_bounds__(256) void k_fwd_last(LsParams p, int LPW, int flags) {
    if (p.ctl->done || (MODE == LS_CG && p.ctl->cg_stop)) return;
    constexpr int NLP = padded<NL>();
    constexpr int D = (C == 3) ? 2 : 3;
    using cx = cplx<SR>;
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* buf = tw + NL;
    const int P = LPW / 2;
    const Geom g = p.g;
    const long line0 = (long)blockIdx.x * LPW;
    stage_twiddles<SR>(tw, reinterpret_cast<const cx*>(MODE == LS_FULL ? p.tw_l64 : p.tw_l32), NL);

    const bool add_shift = (MODE == LS_FULL) && (flags & LS_FLAG_SHIFT);
    const bool add_corr = (MODE == LS_FULL) && (flags & LS_FLAG_CORR) && p.ctl->corr_valid;
    double shift[C];
#pragma unroll
    for (int c = 0; c < C; ++c) shift[c] = add_shift ? p.ctl->shift[c] : 0.0;
    SR* bufr = reinterpret_cast<SR*>(buf);
    for (int vox = threadIdx.x; vox < LPW * NL; vox += blockDim.x) {
        const int li = vox / NL, j = vox % NL;
        const long line = line0 + li;
        SR tau[C];
        if (line < g.nlines()) {
            const long v = line * NL + j;
            if constexpr (MODE == LS_FULL) {
                double e[C], ee[C];
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    e[c] = p.eps[c * g.N + v] + shift[c];
                    if (add_corr) e[c] += (double)p.corr[c * g.N + v];
                    ee[c] = p.epsp ? e[c] - p.epsp[c * g.N + v] : e[c];
                }
                const double mu = p.mu[v], lam = p.lam[v];
                double tre = 0.0, tr0 = 0.0;
#pragma unroll
                for (int c = 0; c < D; ++c) { tre += ee[c]; tr0 += e[c]; }
                const double tm = 2.0 * mu, lt = lam * tre, tm0 = 2.0 * p.mu0, lt0 = p.lam0 * tr0;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const double s = tm * ee[c] + (c < D ? lt : 0.0);
                    const double s0 = tm0 * e[c] + (c < D ? lt0 : 0.0);
                    tau[c] = (SR)(s - s0);
                }
            } else {
                // CG vector updates (Chronopoulos-Gear, lines 7-10), then tau = (C - C0):r
                float y[C];
                const bool upd = flags & (LS_FLAG_CG_UPDATE1 | LS_FLAG_CG_UPDATE);
                if (upd) {
                    const float al = (float)p.ctl->cg_alpha, be = (float)p.ctl->cg_beta;
                    const bool first = flags & LS_FLAG_CG_UPDATE1;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const long o = c * g.N + v;
                        const float r = p.r32[o], w = p.w32[o];
                        const float pp = first ? r : r + be * p.p32[o];
                        const float ss = first ? w : w + be * p.s32[o];
                        // x is the correction of eps_1 = eps_0 + r0, the field the FP64
                        // step left: x_j = d_j - r0 where d_1 = alpha_0 r0 solves A d = r0
                        const float xx = first ? (al - 1.0f) * r : p.corr[o] + al * pp;
                        const float rn = r - al * ss;
                        p.p32[o] = pp;
                        p.s32[o] = ss;
                        p.corr[o] = xx;
                        p.r32[o] = rn;
                        y[c] = rn;
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < C; ++c) y[c] = p.r32[c * g.N + v];
                }
                float tr = 0.f;
#pragma unroll
                for (int c = 0; c < D; ++c) tr += y[c];
                const float dm2 = 2.0f * p.dmu32[v], dl = p.dlam32[v] * tr;
#pragma unroll
                for (int c = 0; c < C; ++c) tau[c] = (SR)(dm2 * y[c] + (c < D ? dl : 0.0f));
            }
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) tau[c] = (SR)0;
        }
        const int pr = li >> 1, part = li & 1;
#pragma unroll
        for (int c = 0; c < C; ++c) bufr[2 * ((c * P + pr) * NLP + j) + part] = tau[c];
    }
    __syncthreads();
    lds_fft<NL, -1, SR>(buf, C * P, NLP, 1, tw);
    // split the packed spectrum into the two half spectra
    cx* spec = reinterpret_cast<cx*>(MODE == LS_FULL ? p.spec64 : p.spec32);
    const int nlh = g.nlh;
    for (int idx = threadIdx.x; idx < LPW * nlh * C; idx += blockDim.x) {
        const int c = idx % C, k = (idx / C) % nlh, li = idx / (C * nlh);
        const long line = line0 + li;
        if (line >= g.nlines()) continue;
        const int pr = li >> 1;
        const cx z = buf[(c * P + pr) * NLP + k];
        const cx zm = buf[(c * P + pr) * NLP + (NL - k) % NL];
        cx out;
        if ((li & 1) == 0) out = {(SR)0.5 * (z.x + zm.x), (SR)0.5 * (z.y - zm.y)};
        else out = {(SR)0.5 * (z.y + zm.y), (SR)-0.5 * (z.x - zm.x)};
        spec[(line * nlh + k) * C + c] = out;
    }
}

// ---------------------------------------------------------------------------
// Passes B/D: c2c along M for one f plane, a tile of (l-bin, component)
// columns.
// ---------------------------------------------------------------------------
template <typename SR, int NM, int DIR>
__global__ __launch_bounds__(256) void k_mid(LsParams p, int TK, int ntiles, int mode) {
    if (p.ctl->done || (mode == LS_CG && p.ctl->cg_stop)) return;
    constexpr int NMP = padded<NM>();
    using cx = cplx<SR>;
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* buf = tw + NM;
    const Geom g = p.g;
    stage_twiddles<SR>(tw, reinterpret_cast<const cx*>(mode == LS_FULL ? p.tw_m64 : p.tw_m32), NM);
    const int ncol = g.nlh * g.C;
    const int tile = blockIdx.x % ntiles;
    const long f = blockIdx.x / ntiles;
    const int k0 = tile * TK;
    const int tk = min(TK, ncol - k0);
    cx* base = reinterpret_cast<cx*>(mode == LS_FULL ? p.spec64 : p.spec32) + f * (long)NM * ncol + k0;
    for (int idx = threadIdx.x; idx < NM * TK; idx += blockDim.x) {
        const int kk = idx % TK, m = idx / TK;
        if (kk < tk) buf[kk * NMP + m] = base[(long)m * ncol + kk];
    }
    __syncthreads();
    lds_fft<NM, DIR, SR>(buf, tk, NMP, 1, tw);
    for (int idx = threadIdx.x; idx < NM * TK; idx += blockDim.x) {
        const int kk = idx % TK, m = idx / TK;
        if (kk < tk) base[(long)m * ncol + kk] = buf[kk * NMP + m];
    }
}

// ---------------------------------------------------------------------------
// Pass C: c2c along F, Green operator, inverse c2c along F.  A block owns TC
// consecutive columns of the flattened (m, l-bin) index, all C components.
// ---------------------------------------------------------------------------
template <typename SR, int NF, int C>
__global__ __launch_bounds__(256) void k_first_gamma(LsParams p, int TC, int mode) {
    if (p.ctl->done || (mode == LS_CG && p.ctl->cg_stop)) return;
    constexpr int NFP = padded<NF>();
    constexpr int D = (C == 3) ? 2 : 3;
    using cx = cplx<SR>;
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* buf = tw + NF;
    const Geom g = p.g;
    stage_twiddles<SR>(tw, reinterpret_cast<const cx*>(mode == LS_FULL ? p.tw_f64 : p.tw_f32), NF);
    const long ncols = (long)g.nm * g.nlh;
    const long col0 = (long)blockIdx.x * TC;
    cx* spec = reinterpret_cast<cx*>(mode == LS_FULL ? p.spec64 : p.spec32);
    for (int idx = threadIdx.x; idx < NF * TC * C; idx += blockDim.x) {
        const int c = idx % C, cc = (idx / C) % TC, f = idx / (C * TC);
        const long col = col0 + cc;
        cx v = {(SR)0, (SR)0};
        if (col < ncols) v = spec[((long)f * ncols + col) * C + c];
        buf[(c * TC + cc) * NFP + f] = v;
    }
    __syncthreads();
    lds_fft<NF, -1, SR>(buf, C * TC, NFP, 1, tw);
    const SR a = (SR)(0.5 / p.mu0);
    const SR B = (SR)((p.lam0 + p.mu0) / (p.mu0 * (p.lam0 + 2.0 * p.mu0)));
    const SR scale = (SR)(1.0 / (double)g.N);
    for (int idx = threadIdx.x; idx < NF * TC; idx += blockDim.x) {
        const int cc = idx % TC, f = idx / TC;
        const long col = col0 + cc;
        if (col >= ncols) continue;
        const int km = (int)(col / g.nlh), kl = (int)(col % g.nlh);
        // signed frequencies k/n as fftfreq (linear_elastic_solver.hpp:22-33)
        auto sfreq = [](int k, int n, bool& nyq) -> SR {
            nyq = (n % 2 == 0) && (k == n / 2) && (n > 1);
            const int lim = (n + 1) / 2;
            return (SR)(k < lim ? k : k - n) / (SR)n;
        };
        bool nf_, nm_, nl_;
        SR qa[3];
        qa[0] = sfreq(f, NF, nf_);
        if (D == 3) {
            qa[1] = sfreq(km, g.nm, nm_);
            qa[2] = sfreq(kl, g.nl, nl_);
        } else {
            nm_ = false;
            qa[1] = sfreq(kl, g.nl, nl_);
        }
        cx t[C], o[C];
#pragma unroll
        for (int c = 0; c < C; ++c) t[c] = buf[(c * TC + cc) * NFP + f];
        SR k2 = (SR)0;
#pragma unroll
        for (int d = 0; d < D; ++d) k2 += qa[d] * qa[d];
        if (k2 == (SR)0) {
#pragma unroll
            for (int c = 0; c < C; ++c) o[c] = {(SR)0, (SR)0};
        } else {
            const SR inv = (SR)1 / sqrt(k2);
            SR q[3];
#pragma unroll
            for (int d = 0; d < D; ++d) q[d] = qa[d] * inv;
            gamma_apply<D, SR>(t, o, q, a, B);
            const bool fl[3] = {nf_, D == 3 ? nm_ : nl_, nl_};
            if (fl[0] || fl[1] || (D == 3 && fl[2])) {
                SR q2[3];
#pragma unroll
                for (int d = 0; d < D; ++d) q2[d] = fl[d] ? -q[d] : q[d];
                cx o2[C];
                gamma_apply<D, SR>(t, o2, q2, a, B);
#pragma unroll
                for (int c = 0; c < C; ++c) o[c] = {(SR)0.5 * (o[c].x + o2[c].x), (SR)0.5 * (o[c].y + o2[c].y)};
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) buf[(c * TC + cc) * NFP + f] = {-o[c].x * scale, -o[c].y * scale};
    }
    __syncthreads();
    lds_fft<NF, +1, SR>(buf, C * TC, NFP, 1, tw);
    for (int idx = threadIdx.x; idx < NF * TC * C; idx += blockDim.x) {
        const int c = idx % C, cc = (idx / C) % TC, f = idx / (C * TC);
        const long col = col0 + cc;
        if (col < ncols) spec[((long)f * ncols + col) * C + c] = buf[(c * TC + cc) * NFP + f];
    }
}

// ---------------------------------------------------------------------------
// Pass E: c2r along L + update.  FULL: norms and the last-block convergence
// test.  DELTA: y <- Gamma~(y), corr (+)= y; no reduction.
// ---------------------------------------------------------------------------
template <typename SR, int NL, int C, int MODE>
__global__ __launch_bounds__(256) void k_inv_last(LsParams p, int LPW, int flags) {
    if (p.ctl->done || (MODE == LS_CG && p.ctl->cg_stop)) return;
    constexpr int NLP = padded<NL>();
    constexpr int D = (C == 3) ? 2 : 3;
    using cx = cplx<SR>;
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* buf = tw + NL;
    __shared__ double red[8 * 4 * 2];
    const int P = LPW / 2;
    const Geom g = p.g;
    const long line0 = (long)blockIdx.x * LPW;
    stage_twiddles<SR>(tw, reinterpret_cast<const cx*>(MODE == LS_FULL ? p.tw_l64 : p.tw_l32), NL);
    const cx* spec = reinterpret_cast<const cx*>(MODE == LS_FULL ? p.spec64 : p.spec32);
    const int nlh = g.nlh;
    for (int idx = threadIdx.x; idx < P * nlh * C; idx += blockDim.x) {
        const int c = idx % C, k = (idx / C) % nlh, pr = idx / (C * nlh);
        const long la = line0 + 2 * pr, lb = la + 1;
        cx A = {(SR)0, (SR)0}, Bv = {(SR)0, (SR)0};
        if (la < g.nlines()) A = spec[(la * nlh + k) * C + c];
        if (lb < g.nlines()) Bv = spec[(lb * nlh + k) * C + c];
        cx* line = buf + (c * P + pr) * NLP;
        if (k == 0 || 2 * k == NL) {
            line[k] = {A.x, Bv.x};   // self-conjugate bins: real parts only (c2r)
        } else {
            line[k] = {A.x - Bv.y, A.y + Bv.x};
            line[NL - k] = {A.x + Bv.y, Bv.x - A.y};
        }
    }
    __syncthreads();
    lds_fft<NL, +1, SR>(buf, C * P, NLP, 1, tw);

    const bool add_shift = (MODE == LS_FULL) && (flags & LS_FLAG_SHIFT);
    const bool add_corr = (MODE == LS_FULL) && (flags & LS_FLAG_CORR) && p.ctl->corr_valid;
    double shift[C], bar[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        shift[c] = add_shift ? p.ctl->shift[c] : 0.0;
        bar[c] = (MODE == LS_FULL) ? p.ctl->eps_bar[c] : 0.0;
    }
    const double mu2 = 2.0 * p.mu0, lam0 = p.lam0;
    const SR* bufr = reinterpret_cast<const SR*>(buf);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int vox = threadIdx.x; vox < LPW * NL; vox += blockDim.x) {
        const int li = vox / NL, j = vox % NL;
        const long line = line0 + li;
        if (line >= g.nlines()) continue;
        const long v = line * NL + j;
        const int pr = li >> 1, part = li & 1;
        if constexpr (MODE == LS_FULL) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const SR et = bufr[2 * ((c * P + pr) * NLP + j) + part];
                const double w = (c < D) ? 1.0 : 2.0;   // off-diagonal entries appear twice in the full tensor
                double old = p.eps[c * g.N + v] + shift[c];
                if (add_corr) old += (double)p.corr[c * g.N + v];
                const double nw = bar[c] + (double)et;
                const double d = nw - old;
                acc[0] += w * d * d;
                acc[1] += w * old * old;
                acc[2] += w * nw * nw;
                p.eps[c * g.N + v] = nw;
                if (p.r32) p.r32[c * g.N + v] = (float)d;
            }
        } else {
            // w = A r = r - Gamma~(r); C0-energy products (r, r), (w, r) and ||r||_W^2
            double rr = 0.0, wr = 0.0, nr = 0.0, trr = 0.0, trw = 0.0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const float et = (float)bufr[2 * ((c * P + pr) * NLP + j) + part];
                const float r = p.r32[c * g.N + v];
                const float w = r - et;
                p.w32[c * g.N + v] = w;
                const double wt = (c < D) ? 1.0 : 2.0;
                rr += wt * (double)r * (double)r;
                wr += wt * (double)w * (double)r;
                if (c < D) { trr += r; trw += w; }
            }
            nr = rr;
            acc[0] += mu2 * rr + lam0 * trr * trr;
            acc[1] += mu2 * wr + lam0 * trw * trr;
            acc[2] += nr;
        }
    }
    block_sum<3>(acc, red);
    if (threadIdx.x == 0) {
        p.partials[3 * blockIdx.x] = acc[0];
        p.partials[3 * blockIdx.x + 1] = acc[1];
        p.partials[3 * blockIdx.x + 2] = acc[2];
    }
}

// Reduction of pass E's block partials and the control-block update.  A kernel
// of its own: the kernel boundary makes every block's partials visible, with
// no per-block device-scope fence (which writes back the XCD's L2).
template <int MODE>
__global__ __launch_bounds__(256) void k_ls_reduce(LsParams p, int nblocks, int flags) {
    if (p.ctl->done || (MODE == LS_CG && p.ctl->cg_stop)) return;
    __shared__ double red[8 * 4 * 2];
    double tot[3];
    reduce_partials<3>(p.partials, nblocks, tot, red);
    if (threadIdx.x == 0) {
        Ctl* ctl = p.ctl;
        if constexpr (MODE == LS_FULL) {
            const double diff = sqrt(tot[0]) / (sqrt(tot[1] + ctl->norm_extra) + 1e-20);
            ctl->diff = diff;
            ctl->diff_sq = tot[0];
            ctl->norm_sq = tot[1];
            ctl->eps_norm = sqrt(tot[2] + ctl->norm_extra);
            if (ctl->iters + ctl->inner < 32) ctl->diff_hist[ctl->iters + ctl->inner] = diff;
            ctl->iters += 1;
            ctl->corr_valid = 0;   // a pending correction was absorbed into eps
            ctl->cg_stop = 0;
            if (diff < ctl->tol || ctl->iters + ctl->inner >= ctl->max_iter) ctl->done = 1;
        } else {
            // Chronopoulos-Gear scalars (alpha, beta) for the next update, and the
            // stopping test on ||r||_W / ||eps||, the relative increment the
            // reference's basic scheme would measure at this iterate.
            const double gam = tot[0], del = tot[1];
            const bool first = flags & LS_FLAG_CG_FIRST;
            ctl->inner += 1;
            if (!first) {
                ctl->corr_valid = 1;
                const double diff = sqrt(tot[2]) / (ctl->eps_norm + 1e-20);
                ctl->diff = diff;
                if (ctl->iters + ctl->inner < 32) ctl->diff_hist[ctl->iters + ctl->inner] = diff;
                if (diff < ctl->tol || ctl->iters + ctl->inner >= ctl->max_iter) ctl->done = 1;
            }
            double al, be;
            if (first) {
                be = 0.0;
                al = gam / del;
            } else {
                be = gam / ctl->cg_gamma;
                al = gam / (del - be * gam / ctl->cg_alpha);
            }
            if (!(al > 0.0) || !isfinite(al) || !isfinite(be)) ctl->cg_stop = 1;   // breakdown: back to FP64
            ctl->cg_alpha = al;
            ctl->cg_beta = be;
            ctl->cg_gamma = gam;
        }
    }
}


// ---------------------------------------------------------------------------
// Field mean -> warm-start shift (linear_elastic_solver.hpp:181-191); or fill
// with eps_bar for a cold start (:193).
// ---------------------------------------------------------------------------
template <int C>
__global__ __launch_bounds__(256) void k_mean_shift(LsParams p) {
    __shared__ double red[8 * 6];
    const Geom g = p.g;
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < g.N; v += (long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] += p.eps[c * g.N + v];
    block_sum<C>(acc, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < C; ++c) p.partials[C * blockIdx.x + c] = acc[c];
}

template <int C>
__global__ __launch_bounds__(256) void k_mean_reduce(LsParams p, int nblocks) {
    __shared__ double red[8 * 6];
    double tot[C];
    reduce_partials<C>(p.partials, nblocks, tot, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < C; ++c) p.ctl->shift[c] = p.ctl->eps_bar[c] - tot[c] / (double)p.g.N;
    }
}

template <int C>
__global__ __launch_bounds__(256) void k_fill_bar(LsParams p) {
    const Geom g = p.g;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < g.N; v += (long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < C; ++c) p.eps[c * g.N + v] = p.ctl->eps_bar[c];
}

// eps += shift (a warm start with no iteration still shifts, :181-191)
template <int C>
__global__ __launch_bounds__(256) void k_apply_shift(LsParams p) {
    const Geom g = p.g;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < g.N; v += (long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < C; ++c) p.eps[c * g.N + v] += p.ctl->shift[c];
}

// ---------------------------------------------------------------------------
// Final stress and macros (linear_elastic_solver.hpp:299-308, :583-592).
// ---------------------------------------------------------------------------
template <int C>
__global__ __launch_bounds__(256) void k_final_stress(LsParams p) {
    constexpr int D = (C == 3) ? 2 : 3;
    __shared__ double red[8 * 12];
    const Geom g = p.g;
    double acc[2 * C];
#pragma unroll
    for (int c = 0; c < 2 * C; ++c) acc[c] = 0.0;
    const bool absorb = p.corr && p.ctl->corr_valid;   // the solve ended inside an FP32 CG cycle
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < g.N; v += (long)gridDim.x * blockDim.x) {
        double e[C], ee[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            e[c] = p.eps[c * g.N + v];
            if (absorb) {
                e[c] += (double)p.corr[c * g.N + v];
                p.eps[c * g.N + v] = e[c];
            }
            ee[c] = p.epsp ? e[c] - p.epsp[c * g.N + v] : e[c];
        }
        double tre = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) tre += ee[c];
        const double tm = 2.0 * p.mu[v], lt = p.lam[v] * tre;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const double s = tm * ee[c] + (c < D ? lt : 0.0);
            p.sig[c * g.N + v] = s;
            acc[c] += e[c];
            acc[C + c] += s;
        }
    }
    block_sum<2 * C>(acc, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < 2 * C; ++c) p.partials[2 * C * blockIdx.x + c] = acc[c];
}

template <int C>
__global__ __launch_bounds__(256) void k_final_reduce(LsParams p, int nblocks) {
    __shared__ double red[8 * 12];
    double tot[2 * C];
    reduce_partials<2 * C>(p.partials, nblocks, tot, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            p.ctl->mean_eps[c] = tot[c] / (double)p.g.N;
            p.ctl->mean_sig[c] = tot[C + c] / (double)p.g.N;
        }
        p.ctl->corr_valid = 0;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

#define MGK_LEN_CASES(X) X(4) X(8) X(16) X(32) X(48) X(64) X(96) X(128) X(256)

bool ls_length_supported(int n) {
#define MGK_CASE(NN) if (n == NN) return true;
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    return false;
}

static int blocks_for(long n, int per) { return (int)((n + per - 1) / per); }

// LDS budget per block (bytes); keeps several blocks resident per CU.
static constexpr int kLdsBudget = 48 * 1024;

static int fft_threads(int n) {
#define MGK_CASE(NN) if (n == NN) return FftPlan<NN>::threads;
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    return 1;
}

// Lines per block of the L passes: an even count whose FFT lanes fit one
// 256-lane pass and whose LDS stays within the budget.
template <typename SR> static int lines_per_block(int nl, int C) {
    int nlp = nl + ((nl % 16 == 0) ? 1 : 0);
    int bytes_per_pair = C * nlp * (int)sizeof(cplx<SR>);
    int pairs_lds = (kLdsBudget - nl * (int)sizeof(cplx<SR>)) / bytes_per_pair;
    int pairs_thr = 256 / (C * fft_threads(nl));
    int pairs = std::max(1, std::min(std::min(pairs_lds, pairs_thr), 8));
    return 2 * pairs;
}

// Columns per block of the F pass, same two limits.
template <typename SR> static int cols_per_block(int nf, int C) {
    int nfp = nf + ((nf % 16 == 0) ? 1 : 0);
    int tc_lds = (kLdsBudget - nf * (int)sizeof(cplx<SR>)) / (C * nfp * (int)sizeof(cplx<SR>));
    int tc_thr = 256 / (C * fft_threads(nf));
    return std::max(1, std::min(std::min(tc_lds, tc_thr), 16));
}

template <typename SR> static size_t smem_bytes(int n, int lines, int nlp) {
    return (size_t)(n + (size_t)lines * nlp) * sizeof(cplx<SR>);
}

template <typename SR, int C, int MODE>
static void launch_fwd_last(const LsParams& p, int flags, hipStream_t s) {
    const int nl = p.g.nl;
    const int lpw = lines_per_block<SR>(nl, C);
    const int nb = blocks_for(p.g.nlines(), lpw);
    const int nlp = nl + ((nl % 16 == 0) ? 1 : 0);
    const size_t sm = smem_bytes<SR>(nl, C * lpw / 2, nlp);
#define MGK_CASE(NN) \
    if (nl == NN) { hipLaunchKernelGGL((k_fwd_last<SR, NN, C, MODE>), dim3(nb), dim3(256), sm, s, p, lpw, flags); return; }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the last axis");
}

template <typename SR, int C, int MODE>
static void launch_inv_last(const LsParams& p, int flags, hipStream_t s) {
    const int nl = p.g.nl;
    const int lpw = lines_per_block<SR>(nl, C);
    const int nb = blocks_for(p.g.nlines(), lpw);
    const int nlp = nl + ((nl % 16 == 0) ? 1 : 0);
    const size_t sm = smem_bytes<SR>(nl, C * lpw / 2, nlp);
#define MGK_CASE(NN) \
    if (nl == NN) { \
        hipLaunchKernelGGL((k_inv_last<SR, NN, C, MODE>), dim3(nb), dim3(256), sm, s, p, lpw, flags); \
        hipLaunchKernelGGL((k_ls_reduce<MODE>), dim3(1), dim3(256), 0, s, p, nb, flags); \
        return; \
    }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the last axis");
}

template <typename SR, int DIR>
static void launch_mid(const LsParams& p, hipStream_t s, int mode) {
    const int nm = p.g.nm;
    const int nmp = nm + ((nm % 16 == 0) ? 1 : 0);
    const int ncol = p.g.nlh * p.g.C;
    int tkmax = (kLdsBudget - nm * (int)sizeof(cplx<SR>)) / (nmp * (int)sizeof(cplx<SR>));
    tkmax = std::max(1, std::min(tkmax, std::max(1, 256 / fft_threads(nm))));
    const int ntiles = (ncol + tkmax - 1) / tkmax;
    const int tk = (ncol + ntiles - 1) / ntiles;
    const int nb = p.g.nf * ntiles;
    const size_t sm = smem_bytes<SR>(nm, tk, nmp);
#define MGK_CASE(NN) \
    if (nm == NN) { hipLaunchKernelGGL((k_mid<SR, NN, DIR>), dim3(nb), dim3(256), sm, s, p, tk, ntiles, mode); return; }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the middle axis");
}

template <typename SR, int C>
static void launch_first_gamma(const LsParams& p, hipStream_t s, int mode) {
    const int nf = p.g.nf;
    const int tc = cols_per_block<SR>(nf, C);
    const long ncols = (long)p.g.nm * p.g.nlh;
    const int nb = blocks_for(ncols, tc);
    const int nfp = nf + ((nf % 16 == 0) ? 1 : 0);
    const size_t sm = smem_bytes<SR>(nf, C * tc, nfp);
#define MGK_CASE(NN) \
    if (nf == NN) { hipLaunchKernelGGL((k_first_gamma<SR, NN, C>), dim3(nb), dim3(256), sm, s, p, tc, mode); return; }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the first axis");
}

int ls_partial_slots(const Geom& g) {
    // largest grid of any reducing kernel times its partial width
    long lines = g.nlines();
    long nb = (lines + 1) / 2;
    return (int)std::max(nb * 3, 1024L * 12);
}

template <int C> static void iteration_full(const LsParams& p, int flags, hipStream_t s) {
    launch_fwd_last<double, C, LS_FULL>(p, flags, s);
    if (C == 6) launch_mid<double, -1>(p, s, LS_FULL);
    launch_first_gamma<double, C>(p, s, LS_FULL);
    if (C == 6) launch_mid<double, +1>(p, s, LS_FULL);
    launch_inv_last<double, C, LS_FULL>(p, flags, s);
}

template <int C> static void iteration_delta(const LsParams& p, int flags, hipStream_t s) {
    launch_fwd_last<float, C, LS_CG>(p, flags, s);
    if (C == 6) launch_mid<float, -1>(p, s, LS_CG);
    launch_first_gamma<float, C>(p, s, LS_CG);
    if (C == 6) launch_mid<float, +1>(p, s, LS_CG);
    launch_inv_last<float, C, LS_CG>(p, flags, s);
}

static int reduce_grid(long n) {
    long nb = (n + 255) / 256;
    return (int)std::min<long>(std::max<long>(nb, 1), 1024);
}

void ls_enqueue_iteration(const LsParams& p, int mode, int flags, hipStream_t s) {
    if (p.g.C == 3) {
        if (mode == LS_FULL) iteration_full<3>(p, flags, s); else iteration_delta<3>(p, flags, s);
    } else {
        if (mode == LS_FULL) iteration_full<6>(p, flags, s); else iteration_delta<6>(p, flags, s);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

void ls_enqueue_pass(const LsParams& p, int pass, int mode, hipStream_t s) {
    const bool full = (mode == LS_FULL);
    if (p.g.C == 3) {
        if (pass == 0) { if (full) launch_fwd_last<double, 3, LS_FULL>(p, 0, s); else launch_fwd_last<float, 3, LS_CG>(p, 0, s); }
        else if (pass == 2) { if (full) launch_first_gamma<double, 3>(p, s, mode); else launch_first_gamma<float, 3>(p, s, mode); }
        else throw ArgError("pass not present in 2D");
    } else {
        if (pass == 0) { if (full) launch_fwd_last<double, 6, LS_FULL>(p, 0, s); else launch_fwd_last<float, 6, LS_CG>(p, 0, s); }
        else if (pass == 1) { if (full) launch_mid<double, -1>(p, s, mode); else launch_mid<float, -1>(p, s, mode); }
        else if (pass == 2) { if (full) launch_first_gamma<double, 6>(p, s, mode); else launch_first_gamma<float, 6>(p, s, mode); }
        else if (pass == 3) { if (full) launch_mid<double, +1>(p, s, mode); else launch_mid<float, +1>(p, s, mode); }
        else throw ArgError("unknown pass");
    }
    MGK_HIP_CHECK(hipGetLastError());
}

void ls_enqueue_start(const LsParams& p, int start, hipStream_t s) {
    const int nb = reduce_grid(p.g.N);
    if (start == LS_START_KEEP) return;
    if (p.g.C == 3) {
        if (start == LS_START_MEAN) {
            hipLaunchKernelGGL((k_mean_shift<3>), dim3(nb), dim3(256), 0, s, p);
            hipLaunchKernelGGL((k_mean_reduce<3>), dim3(1), dim3(256), 0, s, p, nb);
        }
        else if (start == LS_START_FILL) hipLaunchKernelGGL((k_fill_bar<3>), dim3(nb), dim3(256), 0, s, p);
        else hipLaunchKernelGGL((k_apply_shift<3>), dim3(nb), dim3(256), 0, s, p);
    } else {
        if (start == LS_START_MEAN) {
            hipLaunchKernelGGL((k_mean_shift<6>), dim3(nb), dim3(256), 0, s, p);
            hipLaunchKernelGGL((k_mean_reduce<6>), dim3(1), dim3(256), 0, s, p, nb);
        }
        else if (start == LS_START_FILL) hipLaunchKernelGGL((k_fill_bar<6>), dim3(nb), dim3(256), 0, s, p);
        else hipLaunchKernelGGL((k_apply_shift<6>), dim3(nb), dim3(256), 0, s, p);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

void ls_enqueue_final(const LsParams& p, hipStream_t s) {
    const int nb = reduce_grid(p.g.N);
    if (p.g.C == 3) {
        hipLaunchKernelGGL((k_final_stress<3>), dim3(nb), dim3(256), 0, s, p);
        hipLaunchKernelGGL((k_final_reduce<3>), dim3(1), dim3(256), 0, s, p, nb);
    } else {
        hipLaunchKernelGGL((k_final_stress<6>), dim3(nb), dim3(256), 0, s, p);
        hipLaunchKernelGGL((k_final_reduce<6>), dim3(1), dim3(256), 0, s, p, nb);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

}  // namespace mgk
