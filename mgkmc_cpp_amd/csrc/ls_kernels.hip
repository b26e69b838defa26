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
//   E  inv_last   : c2r along L, eps_new = eps_bar + eps_tilde and the
//                   relative-increment norms (block partials)
//   k_ls_reduce   : one block reduces the partials and sets the convergence
//                   flag in the control block
//
// On 3D grids with square nm = nl planes the FP32 CG steps run as plane
// passes instead: k_fwd_plane (A and B for one component of one f-plane),
// pass C on a tiled spectrum, k_inv_plane (D and E); the forward plane pass
// folds in the previous step's reduction.
//
// Every kernel returns at once when the control block says the solve is done,
// so the host enqueues iterations in batches without a sync per iteration.
//
// Spectrum layout: component-interleaved half spectrum,
//   spec[((f*nm + m)*nlh + l)*C + c]   (complex),
// so every pass moves contiguous rows of (bin, component) pairs (the plane
// passes tile it: runs of kPlaneTile bins per component).
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
#include <atomic>
#include <cstdlib>
#include <type_traits>

namespace mgk {

extern __shared__ __align__(16) unsigned char smem_raw[];

// The parameter block of this block row: the kernel argument for a single
// solve (B = 1; kernel arguments are preloaded), P[blockIdx.y] for a batch.
template <bool BATCH, typename T> __device__ __forceinline__ const T& param_sel(const T& a, const T* P) {
    if constexpr (BATCH) return P[blockIdx.y];
    else return a;
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

// A CG step past this solve's cycle length (a batch enqueues its longest
// trajectory's cycle; the others skip the extra steps, as they would not run
// them alone).
// The CG steps of the current refinement cycle: the first cycle (after the
// first FP64 iteration) and the later ones have lengths of their own.
__device__ __forceinline__ int cycle_len_now(const Ctl* ctl) {
    return ctl->iters <= 1 ? ctl->cycle_len : ctl->cycle_len2;
}
__device__ __forceinline__ bool past_cycle(const Ctl* ctl, int step) {
    const int len = cycle_len_now(ctl);
    return len > 0 && step >= len;
}

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

// Fixed-order reduction of part[b*K + k] over blocks by the first VT lanes of
// one block (the others add zeros), loads through ld(i).  With VT = 256 a
// wider block sums in exactly the order of a 256-lane one.
template <int K, class Ld>
__device__ __forceinline__ void reduce_partials_ld(const Ld& ld, int nblocks, int VT, double (&out)[K], double* scratch) {
    constexpr int U = 4;   // partial rows per lane loaded together
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    const int b_first = ((int)threadIdx.x < VT) ? (int)threadIdx.x : nblocks;
    for (int b0 = b_first; b0 < nblocks; b0 += U * VT) {
        double t[U][K];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * VT;
            const long bb = (b < nblocks) ? b : 0;
#pragma unroll
            for (int k = 0; k < K; ++k) t[u][k] = ld(bb * K + k);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b0 + u * VT < nblocks)
#pragma unroll
                for (int k = 0; k < K; ++k) v[k] += t[u][k];
    }
    block_sum<K>(v, scratch);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = v[k];
}

// The same with plain loads over the whole block (the partials were written
// by an earlier kernel on the same stream).
template <int K> __device__ __forceinline__ void reduce_partials(const double* part, int nblocks, double (&out)[K],
                                                                 double* scratch) {
    reduce_partials_ld<K>([part](long i) { return part[i]; }, nblocks, (int)blockDim.x, out, scratch);
}

// ---------------------------------------------------------------------------
// Device-coherent accesses and the grid barrier of a persistent launch (all
// blocks resident at once).
// The blocks of one persistent launch exchange the spectrum and the reduction
// partials through buffer loads and stores with the sc1 cache-policy bit
// (device scope): they bypass the XCD-private L2 caches, so a grid barrier
// needs no release/acquire fence, which on MI355X writes back and invalidates
// the L2 and costs several microseconds per barrier.  Buffer accesses are
// also bounds-checked by the hardware against the resource's size.
// ---------------------------------------------------------------------------
constexpr int kCohPolicy = 16;                 // the sc1 bit of the buffer instructions' cache policy
constexpr int kBufferWord3 = 0x00020000;       // raw (untyped, unswizzled) buffer resource on gfx9

struct CohBuf {
    __amdgpu_buffer_rsrc_t rs;
    __device__ CohBuf(const void* base, unsigned bytes)
        : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kBufferWord3)) {}
    __device__ __forceinline__ cplx<float> ldc(long i) const {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)(i * 8), 0, kCohPolicy);
        return __builtin_bit_cast(cplx<float>, v);
    }
    __device__ __forceinline__ void stc(long i, cplx<float> x) const {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, 0, 0, 0);   // only for its type
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(v), x), rs, (int)(i * 8), 0, kCohPolicy);
    }
    __device__ __forceinline__ double ldd(long i) const {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)(i * 8), 0, kCohPolicy));
    }
    __device__ __forceinline__ void std_(long i, double x) const {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, 0, 0, 0);   // only for its type
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(v), x), rs, (int)(i * 8), 0, kCohPolicy);
    }
};

// A strided line of complex FP32 values through a CohBuf (the column
// accessor of the F pass inside the persistent launch).
struct CohLine {
    CohBuf b;
    long off, es;
    __device__ __forceinline__ cplx<float> operator()(int j) const { return b.ldc(off + j * es); }
    __device__ __forceinline__ void operator()(int j, cplx<float> v) const { b.stc(off + j * es, v); }
};

// Grid barrier words (unsigned, one 128-byte line each): [0] groups arrived,
// [32 (1 + x)] blocks of group x arrived, [32 (9 + x)] release epoch of group
// x, [32 * 17] error flag.  Group x holds the blocks with blockIdx.x % 8 == x
// (the XCD the block runs on), so at most 8 counters and 8 flags are shared
// across XCDs.  The words are zeroed before every launch that uses them.
constexpr int kBarWords = 32 * 18;
constexpr unsigned kSpinLimit = 1u << 21;      // polls before a barrier gives up (seconds)

// Every block of the launch calls this with the same epoch (1, 2, ...).  A
// lane's stores are performed before its block arrives (s_waitcnt vmcnt(0)),
// and the data exchanged across the barrier moves through CohBuf accesses.
// Lane 0 arrives with no load ahead of its atomic, and each poll loads the
// group's release word and the error flag together (one round trip per
// poll).  Returns false once the error flag is set, by this block on a
// timeout or by another block, so every block leaves a failed launch.
__device__ __forceinline__ bool grid_sync(unsigned* w, unsigned epoch, int* s_flag) {
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nb = gridDim.x;
        const unsigned ng = nb < 8 ? nb : 8u;
        const unsigned x = blockIdx.x % 8;
        const unsigned nx = nb / 8 + (x < nb % 8 ? 1u : 0u);
        int err = 0;
        const unsigned old = __hip_atomic_fetch_add(w + 32 * (1 + x), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == epoch * nx) {
            const unsigned top = __hip_atomic_fetch_add(w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (top + 1 == epoch * ng)
                for (unsigned q = 0; q < ng; ++q)
                    __hip_atomic_store(w + 32 * (9 + q), epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        for (unsigned spins = 0;; ++spins) {
            const unsigned rel = __hip_atomic_load(w + 32 * (9 + x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned bad = __hip_atomic_load(w + 32 * 17, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (bad != 0 || (rel < epoch && spins >= kSpinLimit)) {
                __hip_atomic_store(w + 32 * 17, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                err = 1;
                break;
            }
            if (rel >= epoch) break;
            __builtin_amdgcn_s_sleep(1);
        }
        *s_flag = err;
    }
    __syncthreads();
    return *s_flag == 0;
}

template <typename SR> __device__ __forceinline__ void stage_twiddles(cplx<SR>* dst, const cplx<SR>* src, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = gld(src + i);
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
// Mandel (orthonormal) components of a symmetric 3x3 tensor T, m_i = T : e_i
// with e1 = I/sqrt3, e2 = (xx - yy)/sqrt2, e3 = (xx + yy - 2 zz)/sqrt6,
// e4..e6 = sqrt2 sym(xy), sqrt2 sym(xz), sqrt2 sym(yz).  An isotropic
// stiffness is diagonal there (3 lam + 2 mu on e1, 2 mu on the rest) and the
// Frobenius product is the plain dot product, so the plane passes' CG vectors
// use these components: no component's update needs another's.
// ---------------------------------------------------------------------------
template <typename R> __device__ __forceinline__ void mandel_to_cart(const R* m, R* t) {
    const R a = m[0] * (R)0.57735026918962576451, b = m[1] * (R)0.70710678118654752440,
            c = m[2] * (R)0.40824829046386301637;
    t[0] = (a + b) + c;
    t[1] = (a - b) + c;
    t[2] = a - (R)2 * c;
    t[3] = m[3] * (R)0.70710678118654752440;
    t[4] = m[4] * (R)0.70710678118654752440;
    t[5] = m[5] * (R)0.70710678118654752440;
}
template <typename R> __device__ __forceinline__ void cart_to_mandel(const R* t, R* m) {
    m[0] = ((t[0] + t[1]) + t[2]) * (R)0.57735026918962576451;
    m[1] = (t[0] - t[1]) * (R)0.70710678118654752440;
    m[2] = ((t[0] + t[1]) - (R)2 * t[2]) * (R)0.40824829046386301637;
    m[3] = t[3] * (R)1.41421356237309504880;
    m[4] = t[4] * (R)1.41421356237309504880;
    m[5] = t[5] * (R)1.41421356237309504880;
}
template <typename R> __device__ __forceinline__ void mandel_to_cart_c(const cplx<R>* m, cplx<R>* t) {
    R mx[6], my[6], tx[6], ty[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { mx[c] = m[c].x; my[c] = m[c].y; }
    mandel_to_cart(mx, tx);
    mandel_to_cart(my, ty);
#pragma unroll
    for (int c = 0; c < 6; ++c) t[c] = {tx[c], ty[c]};
}
template <typename R> __device__ __forceinline__ void cart_to_mandel_c(const cplx<R>* t, cplx<R>* m) {
    R tx[6], ty[6], mx[6], my[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { tx[c] = t[c].x; ty[c] = t[c].y; }
    cart_to_mandel(tx, mx);
    cart_to_mandel(ty, my);
#pragma unroll
    for (int c = 0; c < 6; ++c) m[c] = {mx[c], my[c]};
}

// ---------------------------------------------------------------------------
// Block shapes (compile time, so every per-thread loop unrolls and a thread
// issues all of its global loads before it waits on any of them).
// ---------------------------------------------------------------------------
static constexpr int kLdsBudget = 48 * 1024;     // passes A and E
static constexpr int kColLdsBudget = 64 * 1024;  // passes B, C and D

// Block shapes, measured at 64^3 and 128^3 (3D) and 128^2 (2D):
//  * passes A/E: two line pairs per block for pass A everywhere and for pass
//    E in 2D, as many as 256 lanes allow for pass E in 3D;
//  * passes B/D: 512 lanes in FP32 (64 columns of 64), 256 in FP64;
//  * pass C in 3D: point tiles spanning whole 128-byte lines (8 points in
//    FP32 on 512 lanes, 4 in FP64), so neighbouring blocks share no line.
// The caps below are build-time tuning knobs over these rules.
#ifndef MGK_FWD_PAIRS_CAP
#define MGK_FWD_PAIRS_CAP 2
#endif
#ifndef MGK_INV_PAIRS_CAP_2D
#define MGK_INV_PAIRS_CAP_2D 2
#endif
#ifndef MGK_INV_PAIRS_CAP_3D
#define MGK_INV_PAIRS_CAP_3D 8
#endif
#ifndef MGK_MID_COLS_CAP
#define MGK_MID_COLS_CAP 256
#endif
#ifndef MGK_GAMMA_POINTS_CAP
#define MGK_GAMMA_POINTS_CAP 256
#endif

constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }
constexpr int pow2floor(int x) { int p = 1; while (2 * p <= x) p *= 2; return p; }

// Lanes per block of the column passes B, C and D.
template <typename SR, int C> constexpr int col_threads() { return (sizeof(SR) == 4 && C == 6) ? 512 : 256; }
constexpr int kColThreadsMax = 512;
// Points per tile of the plane passes' spectrum layout: point xi of component
// c in plane f at f*nxi*6 + (xi / kPlaneTile)*6*kPlaneTile + c*kPlaneTile +
// xi % kPlaneTile (nxi a multiple of kPlaneTile).
constexpr int kPlaneTile = 8;

// Line pairs per block of the L passes (two real lines share one complex FFT
// line): a power of two whose FFT lanes fit 256 and whose LDS fits the budget.
template <typename SR, int N, int C, int CAP> constexpr int last_pairs() {
    constexpr int T = FftPlan<N>::threads;
    constexpr int by_thr = 256 / (C * T);
    constexpr int by_lds = (int)((kLdsBudget - N * (int)sizeof(cplx<SR>)) / (C * padded<N>() * (int)sizeof(cplx<SR>)));
    return pow2floor(cmax(1, cmin(cmin(by_thr, by_lds), CAP)));
}
template <typename SR, int N, int C> constexpr int fwd_pairs() { return last_pairs<SR, N, C, MGK_FWD_PAIRS_CAP>(); }
// Block-shape variants the autotuner times: HALF = 1 halves pass A's line
// pairs and the M passes' columns per block (more, smaller blocks).  The
// arithmetic of a line or column does not depend on them.
template <typename SR, int N, int C, int HALF> constexpr int fwd_pairs_v() {
    return HALF ? cmax(1, fwd_pairs<SR, N, C>() / 2) : fwd_pairs<SR, N, C>();
}
template <typename SR, int N, int C> constexpr int inv_pairs() {
    return last_pairs<SR, N, C, (C == 6) ? MGK_INV_PAIRS_CAP_3D : MGK_INV_PAIRS_CAP_2D>();
}

// Columns per block of the M passes (3D only, C = 6).
template <typename SR, int N> constexpr int mid_cols() {
    constexpr int T = FftPlan<N>::threads;
    constexpr int by_lds = (int)((kColLdsBudget - N * (int)sizeof(cplx<SR>)) / (padded<N>() * (int)sizeof(cplx<SR>)));
    return pow2floor(cmax(1, cmin(cmin(col_threads<SR, 6>() / T, by_lds), MGK_MID_COLS_CAP)));
}
template <typename SR, int N, int HALF> constexpr int mid_cols_v() {
    return HALF ? cmax(1, mid_cols<SR, N>() / 2) : mid_cols<SR, N>();
}

// Frequency points (all C components each) per block of the F pass with
// LANES lanes per block.
// TILED: the plane passes' spectrum, read in whole tiles of kPlaneTile points.
template <typename SR, int N, int C, int LANES = col_threads<SR, C>(), bool TILED = false> constexpr int gamma_points() {
    constexpr int T = FftPlan<N>::threads;
    constexpr int by_lds =
        (int)((kColLdsBudget - N * (int)sizeof(cplx<SR>)) / (C * padded<N>() * (int)sizeof(cplx<SR>)));
    constexpr int by_thr = LANES / (T * C);
    constexpr int most = cmax(1, cmin(cmin(by_thr, by_lds), MGK_GAMMA_POINTS_CAP));
    constexpr int point_bytes = C * (int)sizeof(cplx<SR>);
    constexpr int g_line = 128 / cgcd(128, point_bytes);   // points per whole number of lines
    constexpr int g_align = TILED ? g_line * kPlaneTile / cgcd(g_line, kPlaneTile) : g_line;
    return (C == 6 && most >= g_align) ? (most / g_align) * g_align : most;
}

// Lanes per block of the F pass: the FFT lanes (G*C lines of T lanes), or
// more when that leaves the Green operator's N*G (point, frequency) items
// short of one per lane; always a multiple of G.
template <typename SR, int N, int C, int LANES = col_threads<SR, C>(), bool TILED = false> constexpr int gamma_threads() {
    constexpr int G = gamma_points<SR, N, C, LANES, TILED>();
    constexpr int fft = G * C * FftPlan<N>::threads;
    constexpr int green = (cmin(N * G, cmax(LANES, kColThreadsMax)) / G) * G;
    return cmax(fft, green);
}

// Lanes of pass C on the plane passes' tiled spectrum: tiles of kPlaneTile
// points need whole tiles per block, which 128-point lines reach with 1024
// lanes (8 points) and not with 512 (5).
template <int NF> constexpr int plane_gamma_lanes() { return NF > 64 ? 1024 : col_threads<float, 6>(); }

// ---------------------------------------------------------------------------
// Pass A: tau + r2c along L.  Lines l = f*nm + m, LPW = 2*PAIRS lines per
// block, VPT voxels per lane.
// ---------------------------------------------------------------------------
template <typename SR, int NL, int C, int MODE, bool BATCH, int HALF = 0>
__global__ __launch_bounds__(256) void k_fwd_last(const LsParams p_arg, const LsParams* __restrict__ P, int flags) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    constexpr int PAIRS = fwd_pairs_v<SR, NL, C, HALF>();
    constexpr int LPW = 2 * PAIRS;
    constexpr int VPT = (LPW * NL + 255) / 256;
    constexpr int NLP = padded<NL>();
    constexpr int NLH = NL / 2 + 1;
    constexpr int D = (C == 3) ? 2 : 3;
    using cx = cplx<SR>;
    const bool stop = (p.ctl->done | (MODE == LS_CG ? (p.ctl->cg_stop | past_cycle(p.ctl, flags >> LS_STEP_SHIFT)) : 0)) != 0;
    if ((flags & LS_FLAG_SPEC) && stop) return;   // a speculative step: skip the loads
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* buf = tw + NL;
    const Geom g = p.g;
    const long nlines = g.nlines();
    const long line0 = (long)blockIdx.x * LPW;

    // gather every per-voxel input of this lane first
    long vid[VPT];
    bool ok[VPT];
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const int vox = threadIdx.x + 256 * k;
        const long line = line0 + vox / NL;
        ok[k] = (vox < LPW * NL) && (line < nlines);
        vid[k] = ok[k] ? line * NL + vox % NL : 0;
    }
    const bool want_corr = (MODE == LS_FULL) && (flags & LS_FLAG_CORR) && p.corr;
    const bool has_epsp = p.epsp != nullptr;
    const bool upd = (MODE == LS_CG) && (flags & (LS_FLAG_CG_UPDATE1 | LS_FLAG_CG_UPDATE));
    const bool upd1 = (MODE == LS_CG) && (flags & LS_FLAG_CG_UPDATE1);
    // Loads are unconditional (lanes past the end read voxel 0; absent
    // optional fields read a valid stand-in) and nothing consumes them before
    // the last one is issued.
    // FULL: e = eps, ep = eps_p, cr = pending FP32 correction.
    // CG:   e = r; w, p, s, x (x in cr).
    using VR = typename std::conditional<MODE == LS_FULL, double, float>::type;
    VR e[VPT][C], ep[VPT][C];
    float cr[VPT][C], w4[VPT][C], p4[VPT][C], s4[VPT][C];
    float mu_[VPT], lam_[VPT];
    double mu64[VPT], lam64[VPT];
    if constexpr (MODE == LS_FULL) {
        const double* src_ep = has_epsp ? p.epsp : p.eps;
        const float* src_cr = want_corr ? p.corr : reinterpret_cast<const float*>(p.eps);
        // the last CG residual r goes with the correction x (w4 holds it here)
        const float* src_r = (want_corr && p.r32) ? p.r32 : reinterpret_cast<const float*>(p.eps);
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                e[k][c] = p.eps[c * g.N + vid[k]];
                ep[k][c] = src_ep[c * g.N + vid[k]];
                cr[k][c] = src_cr[c * g.N + vid[k]];
                w4[k][c] = src_r[c * g.N + vid[k]];
            }
            mu64[k] = p.mu[vid[k]];
            lam64[k] = p.lam[vid[k]];
        }
    } else {
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
#pragma unroll
            for (int c = 0; c < C; ++c) e[k][c] = p.r32[c * g.N + vid[k]];
            mu_[k] = p.dmu32[vid[k]];
            lam_[k] = p.dlam32[vid[k]];
        }
        if (upd) {
#pragma unroll
            for (int k = 0; k < VPT; ++k)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const long o = c * g.N + vid[k];
                    w4[k][c] = p.w32[o];
                    cr[k][c] = p.corr[o];
                    p4[k][c] = p.p32[o];
                    s4[k][c] = p.s32[o];
                }
        }
    }
    stage_twiddles<SR>(tw, reinterpret_cast<const cx*>(MODE == LS_FULL ? p.tw_l64 : p.tw_l32), NL);
    if (stop) return;

    SR* bufr = reinterpret_cast<SR*>(buf);
    if constexpr (MODE == LS_FULL) {
        const bool add_shift = flags & LS_FLAG_SHIFT;
        const bool add_corr = want_corr && p.ctl->corr_valid;
        double shift[C];
#pragma unroll
        for (int c = 0; c < C; ++c) shift[c] = add_shift ? p.ctl->shift[c] : 0.0;
        const double tm0 = 2.0 * p.mu0;
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
            const int vox = threadIdx.x + 256 * k;
            if (vox >= LPW * NL) continue;
            const int li = vox / NL, j = vox % NL;
            double ee[C], ek[C];
            double tre = 0.0, tr0 = 0.0;
            // a pending CG cycle's iterate: x, then its last residual r (the
            // increment the basic scheme would add there), as the final pass
            // absorbs them (shift is 0 whenever a correction is pending)
            double crc[C], rrc[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                crc[c] = (double)cr[k][c];
                rrc[c] = p.r32 ? (double)w4[k][c] : 0.0;
            }
            if constexpr (C == 6) {
                if (add_corr && p.mandel) {   // the CG vectors in Mandel components
                    double t[6];
                    mandel_to_cart(crc, t);
#pragma unroll
                    for (int c = 0; c < 6; ++c) crc[c] = t[c];
                    mandel_to_cart(rrc, t);
#pragma unroll
                    for (int c = 0; c < 6; ++c) rrc[c] = t[c];
                }
            }
#pragma unroll
            for (int c = 0; c < C; ++c) {
                ek[c] = e[k][c] + shift[c];
                if (add_corr) {
                    ek[c] += crc[c];
                    ek[c] += rrc[c];
                }
                ee[c] = has_epsp ? ek[c] - ep[k][c] : ek[c];
            }
#pragma unroll
            for (int c = 0; c < D; ++c) { tre += ee[c]; tr0 += ek[c]; }
            const double tm = 2.0 * mu64[k], lt = lam64[k] * tre, lt0 = p.lam0 * tr0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const double s = tm * ee[c] + (c < D ? lt : 0.0);
                const double s0 = tm0 * ek[c] + (c < D ? lt0 : 0.0);
                bufr[2 * ((c * PAIRS + (li >> 1)) * NLP + j) + (li & 1)] = ok[k] ? (SR)(s - s0) : (SR)0;
            }
        }
    } else {
        // CG vector updates (Chronopoulos-Gear), then tau = (C - C0):r
        const float al = (float)p.ctl->cg_alpha, be = (float)p.ctl->cg_beta;
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
            const int vox = threadIdx.x + 256 * k;
            if (vox >= LPW * NL) continue;
            const int li = vox / NL, j = vox % NL;
            float y[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (upd && ok[k]) {
                    const long o = c * g.N + vid[k];
                    const float r = e[k][c], w = w4[k][c];
                    const float pp = upd1 ? r : r + be * p4[k][c];
                    const float ss = upd1 ? w : w + be * s4[k][c];
                    // x is the correction of eps_1 = eps_0 + r0, the field the FP64
                    // step left: x_j = d_j - r0 where d_1 = alpha_0 r0 solves A d = r0
                    const float xx = upd1 ? (al - 1.0f) * r : cr[k][c] + al * pp;
                    const float rn = r - al * ss;
                    p.p32[o] = pp;
                    p.s32[o] = ss;
                    p.corr[o] = xx;
                    p.r32[o] = rn;
                    y[c] = rn;
                } else {
                    y[c] = e[k][c];
                }
            }
            float tr = 0.f;
#pragma unroll
            for (int c = 0; c < D; ++c) tr += y[c];
            const float dm2 = 2.0f * mu_[k], dl = lam_[k] * tr;
#pragma unroll
            for (int c = 0; c < C; ++c)
                bufr[2 * ((c * PAIRS + (li >> 1)) * NLP + j) + (li & 1)] =
                    ok[k] ? (SR)(dm2 * y[c] + (c < D ? dl : 0.0f)) : (SR)0;
        }
    }
    __syncthreads();
    lds_fft<NL, -1, SR>(buf, C * PAIRS, NLP, 1, tw);
    // split the packed spectrum into the two half spectra
    cx* spec = reinterpret_cast<cx*>(MODE == LS_FULL ? p.spec64 : p.spec32);
    constexpr int SPT = (LPW * NLH * C + 255) / 256;
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int idx = threadIdx.x + 256 * k;
        if (idx >= LPW * NLH * C) continue;
        const int c = idx % C, kb = (idx / C) % NLH, li = idx / (C * NLH);
        const long line = line0 + li;
        if (line >= nlines) continue;
        const int pr = li >> 1;
        const cx z = buf[(c * PAIRS + pr) * NLP + kb];
        const cx zm = buf[(c * PAIRS + pr) * NLP + (NL - kb) % NL];
        cx out;
        if ((li & 1) == 0) out = {(SR)0.5 * (z.x + zm.x), (SR)0.5 * (z.y - zm.y)};
        else out = {(SR)0.5 * (z.y + zm.y), (SR)-0.5 * (z.x - zm.x)};
        spec[(line * NLH + kb) * C + c] = out;
    }
}

// ---------------------------------------------------------------------------
// Passes B/D: c2c along M.  Columns are the (f, l-bin, component) triples,
// contiguous in (l-bin, component) for one f, stride nlh*C along m.  A block
// owns L consecutive columns; lane (l, t) = (tid % L, tid / L), so a wave's
// gather and store touch consecutive columns of one row.
// ---------------------------------------------------------------------------
template <typename SR, int NM, int DIR, bool BATCH, int HALF = 0>
__global__ __launch_bounds__(kColThreadsMax) void k_mid(const LsParams p_arg, const LsParams* __restrict__ P, int mode,
                                                        int speculative, int step) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    constexpr int T = FftPlan<NM>::threads;
    constexpr int L = mid_cols_v<SR, NM, HALF>();
    constexpr int NMP = padded<NM>();
    using cx = cplx<SR>;
    const bool stop = (p.ctl->done | (mode == LS_CG ? (p.ctl->cg_stop | past_cycle(p.ctl, step)) : 0)) != 0;
    if (speculative && stop) return;   // a speculative step: skip the loads
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* work = tw + NM;
    const Geom g = p.g;
    const int ncol = g.nlh * g.C;
    const long total = (long)g.nf * ncol;
    const int l = threadIdx.x % L, t = threadIdx.x / L;
    const long F = (long)blockIdx.x * L + l;
    const bool active = F < total;
    const long f = active ? F / ncol : 0;
    const long col = active ? F % ncol : 0;
    cx* spec = reinterpret_cast<cx*>(mode == LS_FULL ? p.spec64 : p.spec32);
    const GLineRef<SR> column{spec + f * (long)NM * ncol + col, ncol};
    cx v[NM / T];
    line_gather<NM, SR>(column, t, active, v);
    stage_twiddles<SR>(tw, reinterpret_cast<const cx*>(mode == LS_FULL ? p.tw_m64 : p.tw_m32), NM);
    if (stop) return;
    line_finish<NM, DIR, SR, false, false>(v, t, active, column, LineRef<SR>{work + l * NMP, 1}, tw);
}

// ---------------------------------------------------------------------------
// Pass C: c2c along F, Green operator, inverse c2c along F.  Columns are the
// (m, l-bin, component) triples, contiguous for one f (stride nm*nlh*C along
// f).  A block owns G frequency points (m, l-bin) with all C components, i.e.
// L = G*C consecutive columns; lane (l, t) = (tid % L, tid / L).
// ---------------------------------------------------------------------------
// Lane roles of the F pass for a block owning the G points from xi0 on.
// Both spectrum layouts keep the C components of a run of points
// contiguous, so lane l reads column xi0*C + l either way: interleaved, lane l
// is (point l / C, component l % C); plane-tiled (tiles of kPlaneTile points
// per component), lane l is (point (l / (C*T8))*T8 + l % T8, component
// (l % (C*T8)) / T8).
template <typename SR, int NF, int C, bool PLANE, int LANES = col_threads<SR, C>()> struct GammaLanes {
    static constexpr int T = FftPlan<NF>::threads;
    static constexpr int G = gamma_points<SR, NF, C, LANES, PLANE>();
    static constexpr int L = G * C;
    static constexpr int T8 = kPlaneTile;
    int l, t, gl, cl;
    bool active;   // lanes past the FFT lanes serve the Green operator only
    __device__ __forceinline__ GammaLanes(long xi0, long nxi, int tid) {
        l = tid % L;
        t = tid / L;
        gl = PLANE ? (l / (C * T8)) * T8 + l % T8 : l / C;
        cl = PLANE ? (l % (C * T8)) / T8 : l % C;
        active = t < T && xi0 + gl < nxi;
    }
};

// The F pass of one block after its column gather: forward c2c along F into
// the block's LDS lines, eps_hat = -Gamma0_eff(xi):tau_hat / N per point, the
// inverse c2c along F, stored through `column` (global memory or CohLine).
// `tw`: the F twiddles in LDS; `work`: G*C padded lines of LDS.
template <typename SR, int NF, int C, bool PLANE, class Col, int LANES = col_threads<SR, C>()>
__device__ __forceinline__ void gamma_core(cplx<SR> (&v)[NF / FftPlan<NF>::threads],
                                           const GammaLanes<SR, NF, C, PLANE, LANES>& ln,
                                           const Col& column, const cplx<SR>* tw, cplx<SR>* work, long xi0,
                                           const Geom& g, double lam0, double mu0, int tid, int nthr) {
    constexpr int D = (C == 3) ? 2 : 3;
    constexpr int G = gamma_points<SR, NF, C, LANES, PLANE>();
    // the FP32 plane spectra hold Mandel components, the FP64 ones Cartesian
    constexpr bool MANDEL = PLANE && C == 6 && std::is_same<SR, float>::value;
    constexpr int NFP = padded<NF>();
    using cx = cplx<SR>;
    const long nxi = (long)g.nm * g.nlh;
    const int t = ln.t;
    const bool active = ln.active;
    const LineRef<SR> mine{work + (ln.gl * C + ln.cl) * NFP, 1};
    line_finish<NF, -1, SR, false, true>(v, t, active, mine, mine, tw);
    __syncthreads();

    const SR a = (SR)(0.5 / mu0);
    const SR B = (SR)((lam0 + mu0) / (mu0 * (lam0 + 2.0 * mu0)));
    const SR scale = (SR)(1.0 / (double)g.N);
    // signed frequencies k/n as fftfreq (linear_elastic_solver.hpp:22-33)
    auto sfreq = [](int k, int n, bool& nyq) -> SR {
        nyq = (n % 2 == 0) && (k == n / 2) && (n > 1);
        const int lim = (n + 1) / 2;
        return (SR)(k < lim ? k : k - n) / (SR)n;
    };
    // The block has a multiple of G lanes (gamma_threads), so a lane keeps
    // its point gg for every f it visits: the point's M and L frequencies
    // are computed once.
    const int gg = tid % G;
    const long xi = xi0 + gg;
    const int km = xi < nxi ? (int)(xi / g.nlh) : 0, kl = xi < nxi ? (int)(xi % g.nlh) : 0;
    bool pm_nyq, pl_nyq;
    const SR pm = sfreq(km, g.nm, pm_nyq), pl = sfreq(kl, g.nl, pl_nyq);
    for (int f = tid / G; f < NF && xi < nxi; f += nthr / G) {
        bool nf_, nm_, nl_;
        SR qa[3];
        qa[0] = sfreq(f, NF, nf_);
        if (D == 3) {
            qa[1] = pm;
            nm_ = pm_nyq;
            qa[2] = pl;
            nl_ = pl_nyq;
            // FFT axes (F, M, L) to directions (x, y, z); M has length 1
            // (frequency 0, no Nyquist bin) when perm is not PERM_XYZ
            if (g.perm == PERM_XZY) {
                const SR t = qa[1]; qa[1] = qa[2]; qa[2] = t;
                const bool b = nm_; nm_ = nl_; nl_ = b;
            } else if (g.perm == PERM_YXZ) {
                const SR t = qa[0]; qa[0] = qa[1]; qa[1] = t;
                const bool b = nf_; nf_ = nm_; nm_ = b;
            }
        } else {
            nm_ = false;
            qa[1] = pl;
            nl_ = pl_nyq;
        }
        cx tt[C], o[C];
#pragma unroll
        for (int c = 0; c < C; ++c) tt[c] = work[(gg * C + c) * NFP + f];
        if constexpr (MANDEL) {   // tau_hat in Mandel components
            cx t6[6];
            mandel_to_cart_c(tt, t6);
#pragma unroll
            for (int c = 0; c < 6; ++c) tt[c] = t6[c];
        }
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
            gamma_apply<D, SR>(tt, o, q, a, B);
            const bool fl[3] = {nf_, D == 3 ? nm_ : nl_, nl_};
            if (fl[0] || fl[1] || (D == 3 && fl[2])) {
                SR q2[3];
#pragma unroll
                for (int d = 0; d < D; ++d) q2[d] = fl[d] ? -q[d] : q[d];
                cx o2[C];
                gamma_apply<D, SR>(tt, o2, q2, a, B);
#pragma unroll
                for (int c = 0; c < C; ++c) o[c] = {(SR)0.5 * (o[c].x + o2[c].x), (SR)0.5 * (o[c].y + o2[c].y)};
            }
        }
        if constexpr (MANDEL) {
            cx m6[6];
            cart_to_mandel_c(o, m6);
#pragma unroll
            for (int c = 0; c < 6; ++c) o[c] = m6[c];
        }
#pragma unroll
        for (int c = 0; c < C; ++c) work[(gg * C + c) * NFP + f] = {-o[c].x * scale, -o[c].y * scale};
    }
    __syncthreads();
    line_gather<NF, SR>(mine, t, active, v);
    line_finish<NF, +1, SR, true, false>(v, t, active, column, mine, tw);
}

template <typename SR, int NF, int C, bool BATCH, bool PLANE = false, int LANES = col_threads<SR, C>()>
__global__ __launch_bounds__(kColThreadsMax) void k_first_gamma(const LsParams p_arg, const LsParams* __restrict__ P, int mode,
                                                                int speculative, int step) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    constexpr int G = gamma_points<SR, NF, C, LANES, PLANE>();
    using cx = cplx<SR>;
    const bool stop = (p.ctl->done | (mode == LS_CG ? (p.ctl->cg_stop | past_cycle(p.ctl, step)) : 0)) != 0;
    if (speculative && stop) return;   // a speculative step: skip the loads
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* work = tw + NF;
    const Geom g = p.g;
    const long nxi = (long)g.nm * g.nlh;
    const long xi0 = (long)blockIdx.x * G;
    cx* spec = reinterpret_cast<cx*>(mode == LS_FULL ? p.spec64 : p.spec32);
    const GammaLanes<SR, NF, C, PLANE, LANES> ln(xi0, nxi, (int)threadIdx.x);
    const GLineRef<SR> column{spec + (ln.active ? xi0 * C + ln.l : 0), nxi * C};
    cx v[NF / FftPlan<NF>::threads];
    line_gather<NF, SR>(column, ln.t, ln.active, v);
    stage_twiddles<SR>(tw, reinterpret_cast<const cx*>(mode == LS_FULL ? p.tw_f64 : p.tw_f32), NF);
    if (stop) return;
    gamma_core<SR, NF, C, PLANE, GLineRef<SR>, LANES>(v, ln, column, tw, work, xi0, g, p.lam0, p.mu0, (int)threadIdx.x,
                                                     (int)blockDim.x);
}

// Pass C on the tiled spectrum of 128-point plane grids: 1024 lanes, so a
// block owns whole tiles (FP32 CG steps only).
template <int NF, bool BATCH>
__global__ __launch_bounds__(1024) void k_first_gamma_wide(const LsParams p_arg, const LsParams* __restrict__ P,
                                                           int speculative, int step) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    constexpr int LANES = plane_gamma_lanes<NF>();
    constexpr int G = gamma_points<float, NF, 6, LANES>();
    using cx = cplx<float>;
    const bool stop = (p.ctl->done | p.ctl->cg_stop | past_cycle(p.ctl, step)) != 0;
    if (speculative && stop) return;
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* work = tw + NF;
    const Geom g = p.g;
    const long nxi = (long)g.nm * g.nlh;
    const long xi0 = (long)blockIdx.x * G;
    cx* spec = reinterpret_cast<cx*>(p.spec32);
    const GammaLanes<float, NF, 6, true, LANES> ln(xi0, nxi, (int)threadIdx.x);
    const GLineRef<float> column{spec + (ln.active ? xi0 * 6 + ln.l : 0), nxi * 6};
    cx v[NF / FftPlan<NF>::threads];
    line_gather<NF, float>(column, ln.t, ln.active, v);
    stage_twiddles<float>(tw, reinterpret_cast<const cx*>(p.tw_f32), NF);
    if (stop) return;
    gamma_core<float, NF, 6, true, GLineRef<float>, LANES>(v, ln, column, tw, work, xi0, g, p.lam0, p.mu0,
                                                          (int)threadIdx.x, (int)blockDim.x);
}

// ---------------------------------------------------------------------------
// Pass E: c2r along L + update.  FULL: eps_new = eps_bar + eps_tilde and the
// norms of the relative-increment test.  CG: w = A r and its C0 products.
// Block partials go to p.partials; k_ls_reduce finishes.
// ---------------------------------------------------------------------------
template <typename SR, int NL, int C, int MODE, bool BATCH>
__global__ __launch_bounds__(256) void k_inv_last(const LsParams p_arg, const LsParams* __restrict__ P, int flags) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    constexpr int PAIRS = inv_pairs<SR, NL, C>();
    constexpr int LPW = 2 * PAIRS;
    constexpr int VPT = (LPW * NL + 255) / 256;
    constexpr int NLP = padded<NL>();
    constexpr int NLH = NL / 2 + 1;
    constexpr int SPT = (PAIRS * NLH * C + 255) / 256;
    constexpr int D = (C == 3) ? 2 : 3;
    using cx = cplx<SR>;
    const bool stop = (p.ctl->done | (MODE == LS_CG ? (p.ctl->cg_stop | past_cycle(p.ctl, flags >> LS_STEP_SHIFT)) : 0)) != 0;
    if ((flags & LS_FLAG_SPEC) && stop) return;
    cx* tw = reinterpret_cast<cx*>(smem_raw);
    cx* buf = tw + NL;
    __shared__ double red[8 * 4 * 2];
    const Geom g = p.g;
    const long nlines = g.nlines();
    const long line0 = (long)blockIdx.x * LPW;

    // gather: the two half spectra of each line pair, and the voxel inputs
    const cx* spec = reinterpret_cast<const cx*>(MODE == LS_FULL ? p.spec64 : p.spec32);
    cx A[SPT], Bv[SPT];
    bool okA[SPT], okB[SPT];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int idx = threadIdx.x + 256 * k;
        const int c = idx % C, kb = (idx / C) % NLH, pr = idx / (C * NLH);
        const long la = line0 + 2 * pr, lb = la + 1;
        const bool in = idx < PAIRS * NLH * C;
        okA[k] = in && la < nlines;
        okB[k] = in && lb < nlines;
        A[k] = spec[okA[k] ? (la * NLH + kb) * C + c : 0];
        Bv[k] = spec[okB[k] ? (lb * NLH + kb) * C + c : 0];
    }
    long vid[VPT];
    bool ok[VPT];
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const int vox = threadIdx.x + 256 * k;
        const long line = line0 + vox / NL;
        ok[k] = (vox < LPW * NL) && (line < nlines);
        vid[k] = ok[k] ? line * NL + vox % NL : 0;
    }
    const bool want_corr = (MODE == LS_FULL) && (flags & LS_FLAG_CORR) && p.corr;
    using VR = typename std::conditional<MODE == LS_FULL, double, float>::type;
    VR old[VPT][C];
    float cr[VPT][C], rr[VPT][C];
    const float* src_cr = want_corr ? p.corr : reinterpret_cast<const float*>(p.eps);
    const float* src_r = (want_corr && p.r32) ? p.r32 : reinterpret_cast<const float*>(p.eps);
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            if constexpr (MODE == LS_FULL) {
                old[k][c] = p.eps[c * g.N + vid[k]];
                cr[k][c] = src_cr[c * g.N + vid[k]];
                rr[k][c] = src_r[c * g.N + vid[k]];
            } else {
                old[k][c] = p.r32[c * g.N + vid[k]];
            }
        }
    }
    stage_twiddles<SR>(tw, reinterpret_cast<const cx*>(MODE == LS_FULL ? p.tw_l64 : p.tw_l32), NL);
    if (stop) return;

#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int idx = threadIdx.x + 256 * k;
        if (idx >= PAIRS * NLH * C) continue;
        const int c = idx % C, kb = (idx / C) % NLH, pr = idx / (C * NLH);
        cx* line = buf + (c * PAIRS + pr) * NLP;
        if (!okA[k]) A[k] = {(SR)0, (SR)0};
        if (!okB[k]) Bv[k] = {(SR)0, (SR)0};
        if (kb == 0 || 2 * kb == NL) {
            line[kb] = {A[k].x, Bv[k].x};   // self-conjugate bins: real parts only (c2r)
        } else {
            line[kb] = {A[k].x - Bv[k].y, A[k].y + Bv[k].x};
            line[NL - kb] = {A[k].x + Bv[k].y, Bv[k].x - A[k].y};
        }
    }
    __syncthreads();
    lds_fft<NL, +1, SR>(buf, C * PAIRS, NLP, 1, tw);

    const double mu2 = 2.0 * p.mu0, lam0 = p.lam0;
    const SR* bufr = reinterpret_cast<const SR*>(buf);
    double acc[3] = {0.0, 0.0, 0.0};
    if constexpr (MODE == LS_FULL) {
        const bool add_shift = flags & LS_FLAG_SHIFT;
        const bool add_corr = want_corr && p.ctl->corr_valid;
        double shift[C], bar[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            shift[c] = add_shift ? p.ctl->shift[c] : 0.0;
            bar[c] = p.ctl->eps_bar[c];
        }
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
            if (!ok[k]) continue;
            const int vox = threadIdx.x + 256 * k;
            const int li = vox / NL, j = vox % NL;
            double crc[C], rrc[C], dd[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                crc[c] = (double)cr[k][c];
                rrc[c] = p.r32 ? (double)rr[k][c] : 0.0;
            }
            if constexpr (C == 6) {
                if (add_corr && p.mandel) {
                    double t[6];
                    mandel_to_cart(crc, t);
#pragma unroll
                    for (int c = 0; c < 6; ++c) crc[c] = t[c];
                    mandel_to_cart(rrc, t);
#pragma unroll
                    for (int c = 0; c < 6; ++c) rrc[c] = t[c];
                }
            }
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const SR et = bufr[2 * ((c * PAIRS + (li >> 1)) * NLP + j) + (li & 1)];
                const double w = (c < D) ? 1.0 : 2.0;   // off-diagonal entries appear twice in the full tensor
                double o = old[k][c] + shift[c];
                if (add_corr) {
                    o += crc[c];
                    o += rrc[c];
                }
                const double nw = bar[c] + (double)et;
                const double d = nw - o;
                acc[0] += w * d * d;
                acc[1] += w * o * o;
                acc[2] += w * nw * nw;
                p.eps[c * g.N + vid[k]] = nw;
                dd[c] = d;
            }
            if (p.r32) {   // the increment starts the CG cycle (Mandel components for the plane passes)
                if constexpr (C == 6) {
                    if (p.mandel) {
                        double m[6];
                        cart_to_mandel(dd, m);
#pragma unroll
                        for (int c = 0; c < 6; ++c) dd[c] = m[c];
                    }
                }
#pragma unroll
                for (int c = 0; c < C; ++c) p.r32[c * g.N + vid[k]] = (float)dd[c];
            }
        }
    } else {
        // w = A r = r - Gamma~(r); C0-energy products (r, r), (w, r) and ||r||_W^2
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
            if (!ok[k]) continue;
            const int vox = threadIdx.x + 256 * k;
            const int li = vox / NL, j = vox % NL;
            double rr = 0.0, wr = 0.0, trr = 0.0, trw = 0.0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const float et = (float)bufr[2 * ((c * PAIRS + (li >> 1)) * NLP + j) + (li & 1)];
                const float r = old[k][c];
                const float w = r - et;
                p.w32[c * g.N + vid[k]] = w;
                const double wt = (c < D) ? 1.0 : 2.0;
                rr += wt * (double)r * (double)r;
                wr += wt * (double)w * (double)r;
                if (c < D) { trr += r; trw += w; }
            }
            acc[0] += mu2 * rr + lam0 * trr * trr;
            acc[1] += mu2 * wr + lam0 * trw * trr;
            acc[2] += rr;
        }
    }
    block_sum<3>(acc, red);
    if (threadIdx.x == 0) {
        gst(p.partials + 3 * blockIdx.x, acc[0]);
        gst(p.partials + 3 * blockIdx.x + 1, acc[1]);
        gst(p.partials + 3 * blockIdx.x + 2, acc[2]);
    }
}

// ---------------------------------------------------------------------------
// The CG reduction of one step from its three partial sums: the stopping test
// on ||r||_W / ||eps|| (the relative increment the reference's basic scheme
// would measure at this iterate) and the Chronopoulos-Gear scalars (alpha,
// beta) of the next update.  `slots`: the step's scalars and count go to slot
// `slot_new` and the previous ones come from the other slot (reductions
// folded into the plane passes), else the single fields.  Only `write`
// updates the control block.
// ---------------------------------------------------------------------------
struct CgDecision {
    double al, be;
    int halt;   // converged, at the iteration cap, or broken down
};

// The FP32 recursive residual of a CG cycle drifts from the true residual by
// rounding in the FP32 operator and vector updates, by up to about kCgDrift
// times the residual the cycle started from (the true residual measured by the
// FP64 iteration after a cycle: at most 1.5e-7 of the cycle's start on 16^3
// trajectories whose flips make up the whole field, 1.1e-7 at 64^3).  The
// recursive residual must reach tol minus that margin, so that the true one is
// below tol.  Where the margin would exceed half the tolerance, a converged
// cycle does not end the solve: it ends the cycle, and the next FP64
// iteration measures the true residual and makes the convergence decision,
// as in the basic scheme.
constexpr double kCgDrift = 3e-7;

// The control-block values a CG reduction reads, the decision it takes.
struct CgIn {
    int inner_prev;        // CG steps counted before this one
    double g_old, a_old;   // the previous step's gamma and alpha (not read by a cycle's first step)
    int iters, max_iter;
    double tol, eps_norm, cycle_r0;
};
struct CgOut {
    double al, be, gam, diff;
    int inner;
    bool done, stop;
};

__device__ __forceinline__ CgIn cg_in(const Ctl* ctl, bool slots, int slot_new) {
    CgIn s;
    s.inner_prev = slots ? ctl->cgs_inner[slot_new ^ 1] : ctl->inner;
    s.g_old = slots ? ctl->cgs_gamma[slot_new ^ 1] : ctl->cg_gamma;
    s.a_old = slots ? ctl->cgs_alpha[slot_new ^ 1] : ctl->cg_alpha;
    s.iters = ctl->iters;
    s.max_iter = ctl->max_iter;
    s.tol = ctl->tol;
    s.eps_norm = ctl->eps_norm;
    s.cycle_r0 = ctl->cycle_r0;
    return s;
}

__device__ __forceinline__ CgOut cg_core(const CgIn& s, const double (&tot)[3], bool first) {
    CgOut o;
    const double gam = tot[0], del = tot[1];
    const int inner = s.inner_prev + 1;
    bool done = false, verify = false;
    double diff = 0.0;
    if (!first) {
        diff = sqrt(tot[2]) / (s.eps_norm + 1e-20);
        const double margin = kCgDrift * s.cycle_r0;
        const bool trusted = margin <= 0.5 * s.tol;
        const bool conv = diff < (trusted ? s.tol - margin : s.tol);
        done = (conv && trusted) || s.iters + inner >= s.max_iter;
        verify = conv && !done;
    }
    double al, be;
    if (first) {
        be = 0.0;
        al = gam / del;
    } else {
        be = gam / s.g_old;
        al = gam / (del - be * gam / s.a_old);
    }
    // breakdown, or a converged cycle to verify: back to FP64
    const bool stop = verify || !(al > 0.0) || !isfinite(al) || !isfinite(be);
    o.al = al;
    o.be = be;
    o.gam = gam;
    o.diff = diff;
    o.inner = inner;
    o.done = done;
    o.stop = stop;
    return o;
}

__device__ __forceinline__ void cg_write(Ctl* ctl, const CgOut& o, bool first, bool slots, int slot_new, int iters) {
    ctl->inner = o.inner;
    if (!first) {
        ctl->corr_valid = 1;
        ctl->diff = o.diff;
        if (iters + o.inner < 32) ctl->diff_hist[iters + o.inner] = o.diff;
        if (o.done) ctl->done = 1;
    }
    if (o.stop) ctl->cg_stop = 1;
    ctl->cg_alpha = o.al;
    ctl->cg_beta = o.be;
    ctl->cg_gamma = o.gam;
    if (slots) {
        ctl->cgs_inner[slot_new] = o.inner;
        ctl->cgs_alpha[slot_new] = o.al;
        ctl->cgs_beta[slot_new] = o.be;
        ctl->cgs_gamma[slot_new] = o.gam;
    }
}

__device__ __forceinline__ CgDecision cg_step_update(Ctl* ctl, const double (&tot)[3], bool first, bool slots,
                                                     int slot_new, bool write) {
    const CgIn s = cg_in(ctl, slots, slot_new);
    const CgOut o = cg_core(s, tot, first);
    if (write) cg_write(ctl, o, first, slots, slot_new, s.iters);
    return {o.al, o.be, (o.done || o.stop) ? 1 : 0};
}

// ---------------------------------------------------------------------------
// Plane passes (3D CG steps on square nm = nl = NP planes).  One block per
// (component c, plane f) does pass A's work for its component and both the L
// and the M transform in LDS, so passes B and D disappear:
//   k_fwd_plane: CG vector update of component c, tau_c = (C - C0):r, r2c
//                along L (two rows packed per complex line), c2c along M;
//   k_inv_plane: c2c along M, c2r along L, w_c = r_c - eps_tilde_c and the
//                block's share of the C0-energy products.
// The CG vectors hold Mandel components (mandel_to_cart), in which
// (C - C0) and C0 are diagonal: a block reads and writes its own component
// only.  The FP64 passes convert where they meet the CG vectors (the
// increment that starts a cycle, the absorbed iterate), pass C around the
// Green operator.  The spectrum is tiled: runs of kPlaneTile points per
// component, the six components' runs side by side, so pass C (PLANE = true)
// reads it as contiguously as the interleaved layout.  Block x is (c, f) =
// (x / nf, x % nf).
// ---------------------------------------------------------------------------
// Lanes per plane block: 256 up to 32-point planes (a 32-point plane's 256
// four-voxel groups one per lane; 512 lanes left half of them idle in the
// vector loops: a 32^3 batch of 46 ran 17.1 k against 22.1 k events/s), else
// 512 (48-point planes: 5586 against 4910 events/s with 576 lanes, one group
// per lane; the persistent cycle's pass C runs on the 512 lanes of its blocks).
template <int NP> constexpr int plane_threads() { return NP <= 32 ? 256 : 512; }

// Four consecutive voxels of one row: the plane passes move the per-voxel
// CG vectors in 16-byte accesses (planes hold a multiple of 64 voxels, and
// every plane starts 16-byte aligned).
typedef float vox4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ vox4 ld_vox4(const float* b, int q) { return gld(reinterpret_cast<const vox4*>(b + 4 * q)); }
__device__ __forceinline__ void st_vox4(float* b, int q, vox4 v) { gst(reinterpret_cast<vox4*>(b + 4 * q), v); }
// the same for vectors a block keeps in LDS
__device__ __forceinline__ vox4 ld_vox4_lds(const float* b, int q) { return *reinterpret_cast<const vox4*>(b + 4 * q); }
__device__ __forceinline__ void st_vox4_lds(float* b, int q, vox4 v) { *reinterpret_cast<vox4*>(b + 4 * q) = v; }

template <int NP> constexpr size_t plane_lds_bytes() {
    return (size_t)(2 * NP + (NP / 2) * padded<NP>() + (NP / 2 + 1) * padded<NP>()) * sizeof(cplx<float>);
}

// Sizes of one (component, plane) block's work.
template <int NP> struct PlaneDims {
    static constexpr int T = plane_threads<NP>();                                // lanes per block
    static constexpr int NL = NP, NM = NP, NLH = NL / 2 + 1, NLP = padded<NL>(), NMP = padded<NM>();
    static constexpr int VOX = NM * NL;
    static constexpr int V4 = VOX / 4;                                          // groups of four voxels
    static constexpr int Q4 = (V4 + T - 1) / T;         // groups per lane
    static constexpr int NS = NM * NLH;                                         // half-spectrum points
    static constexpr int SPT = (NS + T - 1) / T;        // points per lane
    static constexpr int PU = (6 * 256 + T - 1) / T;    // partial rows per lane (nf <= 256)
    static constexpr int QC = Q4 > 2 ? 2 : Q4;                                  // groups a lane holds at once
};

// Offset of plane point idx = m*NLH + kb within its component's runs of the
// tiled spectrum (plane f, component c start at f*NM*NLH*6 + c*kPlaneTile).
__device__ __forceinline__ long tile_off(int idx) {
    return (long)(idx / kPlaneTile) * (6 * kPlaneTile) + idx % kPlaneTile;
}

// (C - C0) of component c in Mandel components: 3 dlam + 2 dmu on e1, 2 dmu
// elsewhere (four voxels).
__device__ __forceinline__ vox4 tau_factor(int c, const vox4& dl, const vox4& dm) {
    vox4 t;
#pragma unroll
    for (int e = 0; e < 4; ++e) t[e] = (c == 0) ? (3.0f * dl[e] + 2.0f * dm[e]) : 2.0f * dm[e];
    return t;
}

// CG update of four voxels of component c (in registers), then
// tau_c = (C - C0)_c r_c (tf: tau_factor) into the packed-row LDS layout (row
// m of the plane is the real or imaginary part of complex line m/2).
template <int NP>
__device__ __forceinline__ void plane_update_tau(int q, bool upd, bool upd1, float al, float be, vox4& r,
                                                 const vox4& w, vox4& pv, vox4& sv, vox4& xv, const vox4& tf,
                                                 float* bufr) {
    using PD = PlaneDims<NP>;
    vox4 y = r;
    if (upd) {
        vox4 pp, ss, xx;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            pp[e] = upd1 ? r[e] : r[e] + be * pv[e];
            ss[e] = upd1 ? w[e] : w[e] + be * sv[e];
            // x is the correction of eps_1 = eps_0 + r0 (see pass A)
            xx[e] = upd1 ? (al - 1.0f) * r[e] : xv[e] + al * pp[e];
            y[e] = r[e] - al * ss[e];
        }
        pv = pp;
        sv = ss;
        xv = xx;
        r = y;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float tau = tf[e] * y[e];
        const int vox = 4 * q + e;
        const int m = vox / PD::NL, j = vox % PD::NL;
        bufr[2 * ((m >> 1) * PD::NLP + j) + (m & 1)] = tau;
    }
}

// Forward in-plane transforms of the packed rows in LDS: r2c along L (two
// rows per complex line, split into the two half spectra), c2c along M; the
// plane's half spectrum leaves through st(idx, value), idx = m*NLH + kb.
template <int NP, class St, typename SR = float>
__device__ __forceinline__ void plane_fwd_fft(cplx<SR>* buf, cplx<SR>* pl, const cplx<SR>* tw_l,
                                              const cplx<SR>* tw_m, const St& st, int tid) {
    using PD = PlaneDims<NP>;
    using cx = cplx<SR>;
    constexpr int NL = PD::NL, NM = PD::NM, NLH = PD::NLH, NLP = PD::NLP, NMP = PD::NMP;
    __syncthreads();
    lds_fft_lane<NL, -1, SR>(buf, NM / 2, NLP, 1, tw_l, tid, PlaneDims<NP>::T);
    // split each packed row pair into the two half spectra, transposed
    for (int idx = tid; idx < (NM / 2) * NLH; idx += PlaneDims<NP>::T) {
        const int pr = idx / NLH, kb = idx % NLH;
        const cx z = buf[pr * NLP + kb];
        const cx zm = buf[pr * NLP + (NL - kb) % NL];
        pl[kb * NMP + 2 * pr] = {(SR)0.5 * (z.x + zm.x), (SR)0.5 * (z.y - zm.y)};
        pl[kb * NMP + 2 * pr + 1] = {(SR)0.5 * (z.y + zm.y), (SR)-0.5 * (z.x - zm.x)};
    }
    __syncthreads();
    lds_fft_lane<NM, -1, SR>(pl, NLH, NMP, 1, tw_m, tid, PlaneDims<NP>::T);
    for (int idx = tid; idx < NM * NLH; idx += PlaneDims<NP>::T) {
        const int m = idx / NLH, kb = idx % NLH;
        st(idx, pl[kb * NMP + m]);
    }
}

// Inverse in-plane transforms: the plane's half spectrum (point
// idx = threadIdx.x + PlaneDims<NP>::T*k in sp[k]) into LDS, c2c along M, c2r
// along L; eps_tilde_c is left in buf's packed-row layout.
template <int NP, typename SR>
__device__ __forceinline__ void plane_inv_fft_staged(cplx<SR>* buf, cplx<SR>* pl, const cplx<SR>* tw_l,
                                                     const cplx<SR>* tw_m, int tid);

template <int NP, typename SR = float>
__device__ __forceinline__ void plane_inv_fft(const cplx<SR> (&sp)[PlaneDims<NP>::SPT], cplx<SR>* buf,
                                              cplx<SR>* pl, const cplx<SR>* tw_l, const cplx<SR>* tw_m, int tid) {
    using PD = PlaneDims<NP>;
    using cx = cplx<SR>;
    constexpr int NL = PD::NL, NM = PD::NM, NLH = PD::NLH, NLP = PD::NLP, NMP = PD::NMP;
#pragma unroll
    for (int k = 0; k < PD::SPT; ++k) {
        const int idx = tid + PlaneDims<NP>::T * k;
        if (idx < NM * NLH) pl[(idx % NLH) * NMP + idx / NLH] = sp[k];
    }
    plane_inv_fft_staged<NP, SR>(buf, pl, tw_l, tw_m, tid);
}

// plane_inv_fft after the spectrum's scatter: bin (m, kb) is in
// pl[kb * NMP + m], not yet behind a barrier.
template <int NP, typename SR>
__device__ __forceinline__ void plane_inv_fft_staged(cplx<SR>* buf, cplx<SR>* pl, const cplx<SR>* tw_l,
                                                     const cplx<SR>* tw_m, int tid) {
    using PD = PlaneDims<NP>;
    using cx = cplx<SR>;
    constexpr int NL = PD::NL, NM = PD::NM, NLH = PD::NLH, NLP = PD::NLP, NMP = PD::NMP;
    __syncthreads();
    lds_fft_lane<NM, +1, SR>(pl, NLH, NMP, 1, tw_m, tid, PlaneDims<NP>::T);
    // pack row pairs into complex lines for the c2r along L
    for (int idx = tid; idx < (NM / 2) * NLH; idx += PlaneDims<NP>::T) {
        const int pr = idx / NLH, kb = idx % NLH;
        const cx A = pl[kb * NMP + 2 * pr], Bv = pl[kb * NMP + 2 * pr + 1];
        cx* line = buf + pr * NLP;
        if (kb == 0 || 2 * kb == NL) {
            line[kb] = {A.x, Bv.x};   // self-conjugate bins: real parts only (c2r)
        } else {
            line[kb] = {A.x - Bv.y, A.y + Bv.x};
            line[NL - kb] = {A.x + Bv.y, Bv.x - A.y};
        }
    }
    __syncthreads();
    lds_fft_lane<NL, +1, SR>(buf, NM / 2, NLP, 1, tw_l, tid, PlaneDims<NP>::T);
}

// w_c = r_c - eps_tilde_c (eps_tilde in buf, as plane_inv_fft leaves it) and
// this lane's share of the C0-energy products (r, r)_C0, (w, r)_C0 (C0 =
// 3 lam0 + 2 mu0 on e1, 2 mu0 elsewhere) and ||r||^2.
template <int NP>
__device__ __forceinline__ void plane_w_partials(const float* bufr, int c, double lam0, double mu0,
                                                 const vox4 (&r)[PlaneDims<NP>::Q4], vox4 (&w)[PlaneDims<NP>::Q4],
                                                 double (&acc)[3], int tid) {
    using PD = PlaneDims<NP>;
    const double c0 = (c == 0) ? 3.0 * lam0 + 2.0 * mu0 : 2.0 * mu0;
    acc[0] = acc[1] = acc[2] = 0.0;
#pragma unroll
    for (int k = 0; k < PD::Q4; ++k) {
        const int q = tid + PlaneDims<NP>::T * k;
        if (q >= PD::V4) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int vox = 4 * q + e;
            const int m = vox / PD::NL, j = vox % PD::NL;
            const float et = bufr[2 * ((m >> 1) * PD::NLP + j) + (m & 1)];
            const float rv = r[k][e];
            w[k][e] = rv - et;
            const double rd = (double)rv, wd = (double)w[k][e];
            acc[0] += rd * rd;
            acc[1] += wd * rd;
            acc[2] += rd * rd;
        }
    }
    acc[0] *= c0;
    acc[1] *= c0;
}

// plane_w_partials for planes whose vector groups a lane cannot hold at once:
// r read and w written in chunks of PlaneDims::QC groups, the partial sums in
// the same order.
template <int NP>
__device__ __forceinline__ void plane_w_partials_stream(const float* bufr, int c, double lam0, double mu0,
                                                        const float* rg, float* wg, double (&acc)[3], int tid) {
    using PD = PlaneDims<NP>;
    constexpr int QC = PD::QC;
    const double c0 = (c == 0) ? 3.0 * lam0 + 2.0 * mu0 : 2.0 * mu0;
    acc[0] = acc[1] = acc[2] = 0.0;
    for (int ch = 0; ch < PD::Q4 / QC; ++ch) {
        vox4 r[QC];
#pragma unroll
        for (int k = 0; k < QC; ++k) {
            const int q = tid + PlaneDims<NP>::T * (ch * QC + k);
            r[k] = ld_vox4(rg, q < PD::V4 ? q : 0);
        }
#pragma unroll
        for (int k = 0; k < QC; ++k) {
            const int q = tid + PlaneDims<NP>::T * (ch * QC + k);
            if (q >= PD::V4) continue;
            vox4 w;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int vox = 4 * q + e;
                const int m = vox / PD::NL, j = vox % PD::NL;
                const float et = bufr[2 * ((m >> 1) * PD::NLP + j) + (m & 1)];
                const float rv = r[k][e];
                w[e] = rv - et;
                const double rd = (double)rv, wd = (double)w[e];
                acc[0] += rd * rd;
                acc[1] += wd * rd;
                acc[2] += rd * rd;
            }
            st_vox4(wg, q, w);
        }
    }
    acc[0] *= c0;
    acc[1] *= c0;
}

// The folded reduction's per-lane sums of the previous step's partial rows
// (row b = threadIdx.x + T*u, loaded into part[.][u]); a
// block_sum of tot follows.
template <int PU, int T>
__device__ __forceinline__ void plane_fold_sum(const double (&part)[3][PU], int nbp, double (&tot)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = 0.0;
#pragma unroll
        for (int u = 0; u < PU; ++u)
            if ((int)threadIdx.x + T * u < nbp) v += part[k][u];
        tot[k] = v;
    }
}

// Mandel component m of four voxels' Cartesian increment (the FP64 plane
// passes' d in the p storage), as the five-pass pass E converts it.
__device__ __forceinline__ vox4 r0_from_cart(const LsParams& p, int m, long base_plane, int q) {
    const long N = p.g.N;
    vox4 out;
    if (m < 3) {
        const vox4 a = ld_vox4(p.p32 + base_plane, q), b = ld_vox4(p.p32 + N + base_plane, q),
                   cz = ld_vox4(p.p32 + 2 * N + base_plane, q);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            double t[6] = {(double)a[e], (double)b[e], (double)cz[e], 0.0, 0.0, 0.0}, mm[6];
            cart_to_mandel(t, mm);
            out[e] = (float)mm[m];
        }
    } else {
        const vox4 a = ld_vox4(p.p32 + (long)m * N + base_plane, q);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, mm[6];
            t[m] = (double)a[e];
            cart_to_mandel(t, mm);
            out[e] = (float)mm[m];
        }
    }
    return out;
}

// PU: partial rows per lane of the folded reduction, 1 where the previous
// step's 6*nf partials fit the block's lanes (the fold then adds the same
// terms in the same order).  Up to 64-point planes at most 128 VGPRs, so two
// blocks share a CU: the batched instantiation took 134 with three rows and
// ran one block per CU.  A 128-point plane's LDS holds one block per CU.
template <int NP, bool BATCH, int PU>
__global__ __launch_bounds__(PlaneDims<NP>::T) __attribute__((amdgpu_waves_per_eu(NP > 64 ? 1 : 4))) void k_fwd_plane(
    const LsParams p_arg, const LsParams* __restrict__ P, int flags) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    using PD = PlaneDims<NP>;
    constexpr int NL = PD::NL, NM = PD::NM, NLP = PD::NLP, VOX = PD::VOX, V4 = PD::V4, Q4 = PD::Q4;
    static_assert(PU >= 1 && PU <= PD::PU, "partial rows per lane");
    using cx = cplx<float>;
    // step == cycle_len: only the folded reduction of the cycle's last step
    const int step = flags >> LS_STEP_SHIFT, cyc = cycle_len_now(p.ctl);
    const bool at_end = cyc > 0 && step == cyc;
    const bool stop = (p.ctl->done | p.ctl->cg_stop | (cyc > 0 && step > cyc)) != 0;
    if ((flags & LS_FLAG_SPEC) && stop) return;
    cx* tw_l = reinterpret_cast<cx*>(smem_raw);
    cx* tw_m = tw_l + NL;
    cx* buf = tw_m + NM;               // NM/2 packed row pairs of NLP
    cx* pl = buf + (NM / 2) * NLP;     // NLH columns of NMP (the half spectrum, transposed)
    const Geom g = p.g;
    const int c = blockIdx.x / g.nf, f = blockIdx.x % g.nf;
    const long base = (long)c * g.N + (long)f * VOX;   // component c of plane f
    const bool upd = flags & (LS_FLAG_CG_UPDATE1 | LS_FLAG_CG_UPDATE);
    const bool upd1 = flags & LS_FLAG_CG_UPDATE1;
    // An updating step folds in the reduction of the step before it: every
    // block reduces the same partials in the same order (so all take the same
    // decision and scalars); block 0 writes the control block, into the slot
    // no block of this launch reads.
    __shared__ double red[16 * 3];
    __shared__ double s_al, s_be;
    __shared__ int s_halt;
    double part[3][PU];
    const int nbp = 6 * g.nf;   // blocks of the previous step's inverse plane pass
    if (upd) {
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            const int b = threadIdx.x + PlaneDims<NP>::T * u;
#pragma unroll
            for (int k = 0; k < 3; ++k) part[k][u] = gld(p.partials + 3 * (b < nbp ? b : 0) + k);
        }
    }

    // The vector groups of a lane are held QC at a time (all of them up to
    // 64-point planes, loaded before the folded reduction; 128-point planes
    // stream them in chunks after it).
    constexpr int QC = PD::QC, NCH = Q4 / QC;
    const vox4 zero4 = {0.f, 0.f, 0.f, 0.f};
    vox4 r[QC], w[QC], pv[QC], sv[QC], xv[QC], dm[QC], dl[QC];
    // the first step of a cycle after an FP64 plane iteration starts from the
    // Cartesian increment (and stores its Mandel components for the steps after)
    const bool r_cart = p.r0_cart && (flags & LS_FLAG_CG_FIRST);
    auto load_chunk = [&](int ch) {
#pragma unroll
        for (int k = 0; k < QC; ++k) {
            const int q = threadIdx.x + PlaneDims<NP>::T * (ch * QC + k);
            const int qq = q < V4 ? q : 0;
            r[k] = r_cart ? r0_from_cart(p, c, (long)f * VOX, qq) : ld_vox4(p.r32 + base, qq);
            dm[k] = ld_vox4(p.dmu32 + (long)f * VOX, qq);
            dl[k] = (c == 0) ? ld_vox4(p.dlam32 + (long)f * VOX, qq) : zero4;
            w[k] = pv[k] = sv[k] = xv[k] = zero4;
            if (upd) {
                w[k] = ld_vox4(p.w32 + base, qq);
                pv[k] = ld_vox4(p.p32 + base, qq);
                sv[k] = ld_vox4(p.s32 + base, qq);
                xv[k] = ld_vox4(p.corr + base, qq);
            }
        }
    };
    load_chunk(0);
    stage_twiddles<float>(tw_l, reinterpret_cast<const cx*>(p.tw_l32), NL);
    stage_twiddles<float>(tw_m, reinterpret_cast<const cx*>(p.tw_m32), NM);
    if (stop) return;

    float al = (float)p.ctl->cg_alpha, be = (float)p.ctl->cg_beta;
    if (upd) {
        double tot[3];
        plane_fold_sum<PU, PD::T>(part, nbp, tot);
        block_sum<3>(tot, red);
        if (threadIdx.x == 0) {
            const bool probe = flags & LS_FLAG_PROBE;
            const int slot_new = (flags & LS_FLAG_ODD) ? 0 : 1;   // the previous step's parity
            const CgDecision d = cg_step_update(p.ctl, tot, upd1, true, slot_new, blockIdx.x == 0 && !probe);
            s_al = d.al;
            s_be = d.be;
            s_halt = probe ? 0 : d.halt;
        }
        __syncthreads();
        if (s_halt || at_end) return;
        al = (float)s_al;
        be = (float)s_be;
    }
    float* bufr = reinterpret_cast<float*>(buf);
    for (int ch = 0; ch < NCH; ++ch) {
        if (ch > 0) load_chunk(ch);
#pragma unroll
        for (int k = 0; k < QC; ++k) {
            const int q = threadIdx.x + PlaneDims<NP>::T * (ch * QC + k);
            if (q >= V4) continue;
            plane_update_tau<NP>(q, upd, upd1, al, be, r[k], w[k], pv[k], sv[k], xv[k], tau_factor(c, dl[k], dm[k]),
                                 bufr);
            if (r_cart) st_vox4(p.r32 + base, q, r[k]);
            if (upd) {
                st_vox4(p.p32 + base, q, pv[k]);
                st_vox4(p.s32 + base, q, sv[k]);
                st_vox4(p.corr + base, q, xv[k]);
                st_vox4(p.r32 + base, q, r[k]);
            }
        }
    }
    cx* spec = reinterpret_cast<cx*>(p.spec32) + (long)f * NM * PD::NLH * 6 + c * kPlaneTile;
    plane_fwd_fft<NP>(buf, pl, tw_l, tw_m, [spec](int idx, cx v) { gst(spec + tile_off(idx), v); }, (int)threadIdx.x);
}

template <int NP, bool BATCH>
__global__ __launch_bounds__(PlaneDims<NP>::T) void k_inv_plane(const LsParams p_arg, const LsParams* __restrict__ P, int flags) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    using PD = PlaneDims<NP>;
    constexpr int NL = PD::NL, NM = PD::NM, NLP = PD::NLP, VOX = PD::VOX, V4 = PD::V4, Q4 = PD::Q4, SPT = PD::SPT;
    using cx = cplx<float>;
    const bool stop = (p.ctl->done | p.ctl->cg_stop | past_cycle(p.ctl, flags >> LS_STEP_SHIFT)) != 0;
    if ((flags & LS_FLAG_SPEC) && stop) return;
    cx* tw_l = reinterpret_cast<cx*>(smem_raw);
    cx* tw_m = tw_l + NL;
    cx* buf = tw_m + NM;
    cx* pl = buf + (NM / 2) * NLP;
    __shared__ double red[16 * 3];
    const Geom g = p.g;
    const int c = blockIdx.x / g.nf, f = blockIdx.x % g.nf;
    const long base = (long)c * g.N + (long)f * VOX;

    const cx* spec = reinterpret_cast<const cx*>(p.spec32) + (long)f * NM * PD::NLH * 6 + c * kPlaneTile;
    cx sp[SPT];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int idx = threadIdx.x + PlaneDims<NP>::T * k;
        sp[k] = gld(spec + tile_off(idx < PD::NS ? idx : 0));
    }
    double acc[3];
    if constexpr (PD::QC == Q4) {
        // every vector group in registers, loaded before the transforms
        vox4 r[Q4], w[Q4];
#pragma unroll
        for (int k = 0; k < Q4; ++k) {
            const int q = threadIdx.x + PlaneDims<NP>::T * k;
            r[k] = ld_vox4(p.r32 + base, q < V4 ? q : 0);
        }
        stage_twiddles<float>(tw_l, reinterpret_cast<const cx*>(p.tw_l32), NL);
        stage_twiddles<float>(tw_m, reinterpret_cast<const cx*>(p.tw_m32), NM);
        if (stop) return;
        plane_inv_fft<NP>(sp, buf, pl, tw_l, tw_m, (int)threadIdx.x);
        plane_w_partials<NP>(reinterpret_cast<const float*>(buf), c, p.lam0, p.mu0, r, w, acc, (int)threadIdx.x);
#pragma unroll
        for (int k = 0; k < Q4; ++k) {
            const int q = threadIdx.x + PlaneDims<NP>::T * k;
            if (q < V4) st_vox4(p.w32 + base, q, w[k]);
        }
    } else {
        // 128-point planes: r streamed in chunks after the transforms
        stage_twiddles<float>(tw_l, reinterpret_cast<const cx*>(p.tw_l32), NL);
        stage_twiddles<float>(tw_m, reinterpret_cast<const cx*>(p.tw_m32), NM);
        if (stop) return;
        plane_inv_fft<NP>(sp, buf, pl, tw_l, tw_m, (int)threadIdx.x);
        const float* bufr = reinterpret_cast<const float*>(buf);
        plane_w_partials_stream<NP>(bufr, c, p.lam0, p.mu0, p.r32 + base, p.w32 + base, acc, (int)threadIdx.x);
    }
    block_sum<3>(acc, red);
    if (threadIdx.x == 0) {
        gst(p.partials + 3 * blockIdx.x, acc[0]);
        gst(p.partials + 3 * blockIdx.x + 1, acc[1]);
        gst(p.partials + 3 * blockIdx.x + 2, acc[2]);
    }
}

// ---------------------------------------------------------------------------
// FP64 plane passes (the basic scheme's iteration on square planes of up to
// 64 points): one block per (Cartesian component c, plane f), three launches
// per iteration instead of five.
//   k_fwd_plane64: tau_c = (C - C0):eps' - C:eps_p of component c (a normal
//                  component's block also reads the other two normal
//                  components for the traces), both in-plane transforms;
//   pass C on the tiled FP64 spectrum (Cartesian components);
//   k_inv_plane64: both inverse transforms, eps_c = eps_bar_c + eps_tilde_c,
//                  the block's share of the relative-increment norms, and in
//                  FP32+refine contexts the increment d_c, written Cartesian
//                  into the p vector's storage: the CG cycle that follows
//                  converts it to Mandel components when it starts
//                  (LsParams::r0_cart).
// eps' = eps + warm-start shift + a pending CG iterate (x, then its last
// residual, Mandel components converted as pass A converts them).
// ---------------------------------------------------------------------------
template <int NP> constexpr size_t plane64_lds_bytes() {
    return (size_t)(2 * NP + (NP / 2) * padded<NP>() + (NP / 2 + 1) * padded<NP>()) * sizeof(cplx<double>);
}

// The FP64 plane passes' voxel loops.  A lane takes its voxels U at a time
// and issues every load of a chunk before it uses any value, so a chunk
// costs one memory round trip; the loop bodies are specialised on the
// block's kind (NORMAL: a normal component, whose stress needs the traces
// over the three normal strains) and on a pending CG iterate (CORR: x, then
// its last residual r, Mandel components, added to eps as pass A adds them).
// A missing eps_p or r reads a valid stand-in buffer and is masked, so the
// loads stay unconditional.  eps' = eps + warm-start shift (+ x + r).
template <int NP> struct Plane64Loop {
    static constexpr int VOX = PlaneDims<NP>::VOX;
    static constexpr int K = (VOX + PlaneDims<NP>::T - 1) / PlaneDims<NP>::T;   // voxels per lane
    static constexpr int U = K < 4 ? K : 4;                               // voxels per chunk
    __device__ static int voxel(int k) { return (int)threadIdx.x + PlaneDims<NP>::T * k; }
};

constexpr double kRsqrt2 = 0.70710678118654752440;

// tau_c = sigma_c(eps' - eps_p) - sigma0_c(eps') of this lane's voxels of
// plane f into the packed-row plane buffer.
template <int NP, bool NORMAL, bool CORR>
__device__ __forceinline__ void plane64_tau(const LsParams& p, int c, int f, const double (&sh)[3], double sh_c,
                                            double* bufr) {
    using L = Plane64Loop<NP>;
    using PD = PlaneDims<NP>;
    constexpr int NL = PD::NL, NLP = PD::NLP, VOX = PD::VOX, K = L::K, U = L::U;
    constexpr int NC = NORMAL ? 3 : 1;
    const long N = p.g.N;
    const bool has_ep = p.epsp != nullptr, has_r = p.r32 != nullptr;
    const double* ep = has_ep ? p.epsp : p.eps;
    const float* rv = has_r ? p.r32 : p.corr;
    const double tm0 = 2.0 * p.mu0, lam0 = p.lam0;
    for (int k0 = 0; k0 < K; k0 += U) {
        double e[NC][U], pe[NC][U], la[U], mu[U];
        float x[NC][U], r[NC][U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = L::voxel(k0 + u);
            const long v = (long)f * VOX + (q < VOX ? q : VOX - 1);
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                const long o = (long)(NORMAL ? j : c) * N + v;
                e[j][u] = gld(p.eps + o);
                pe[j][u] = gld(ep + o);
                if constexpr (CORR) {
                    x[j][u] = gld(p.corr + o);
                    r[j][u] = gld(rv + o);
                }
            }
            la[u] = gld(p.lam + v);
            mu[u] = gld(p.mu + v);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = L::voxel(k0 + u);
            if (k0 + u >= K || q >= VOX) continue;
            double ek_c, ee_c, tr0 = 0.0, tre = 0.0;
            if constexpr (NORMAL) {
                double tc[6] = {0, 0, 0, 0, 0, 0}, tr[6] = {0, 0, 0, 0, 0, 0};
                if constexpr (CORR) {
                    const double crc[6] = {(double)x[0][u], (double)x[1][u], (double)x[2][u], 0.0, 0.0, 0.0};
                    const double rrc[6] = {has_r ? (double)r[0][u] : 0.0, has_r ? (double)r[1][u] : 0.0,
                                           has_r ? (double)r[2][u] : 0.0, 0.0, 0.0, 0.0};
                    mandel_to_cart(crc, tc);
                    mandel_to_cart(rrc, tr);
                }
                double ek[3], ee[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    ek[j] = e[j][u] + sh[j];
                    if constexpr (CORR) {
                        ek[j] += tc[j];
                        ek[j] += tr[j];
                    }
                    ee[j] = has_ep ? ek[j] - pe[j][u] : ek[j];
                }
#pragma unroll
                for (int j = 0; j < 3; ++j) { tre += ee[j]; tr0 += ek[j]; }
                ek_c = (c == 0) ? ek[0] : (c == 1) ? ek[1] : ek[2];
                ee_c = (c == 0) ? ee[0] : (c == 1) ? ee[1] : ee[2];
            } else {
                double ev = e[0][u] + sh_c;
                if constexpr (CORR) {
                    ev += (double)x[0][u] * kRsqrt2;
                    ev += (has_r ? (double)r[0][u] : 0.0) * kRsqrt2;
                }
                ek_c = ev;
                ee_c = has_ep ? ev - pe[0][u] : ev;
            }
            const double tm = 2.0 * mu[u], lt = la[u] * tre, lt0 = lam0 * tr0;
            const double sv = tm * ee_c + (NORMAL ? lt : 0.0);
            const double s0 = tm0 * ek_c + (NORMAL ? lt0 : 0.0);
            const int m = q / NL, jl = q % NL;
            bufr[2 * ((m >> 1) * NLP + jl) + (m & 1)] = sv - s0;
        }
    }
}

// The update of pass E for component c of this lane's voxels of plane f:
// eps_c = eps_bar_c + eps_tilde_c (eps_tilde in the packed-row buffer), its
// shares of ||eps_new - eps'||^2, ||eps'||^2, ||eps_new||^2 (weight wgt), and
// in FP32+refine contexts the increment in FP32 (the next CG cycle's start).
template <int NP, bool NORMAL, bool CORR>
__device__ __forceinline__ void plane64_update(const LsParams& p, int c, int f, double sh_c, double bar, double wgt,
                                               const double* bufr, double (&acc)[3]) {
    using L = Plane64Loop<NP>;
    using PD = PlaneDims<NP>;
    constexpr int NL = PD::NL, NLP = PD::NLP, VOX = PD::VOX, K = L::K, U = L::U;
    constexpr int NC = NORMAL ? 3 : 1;
    const long N = p.g.N;
    const bool has_r = p.r32 != nullptr;
    const float* rv = has_r ? p.r32 : p.corr;
    for (int k0 = 0; k0 < K; k0 += U) {
        double e[U];
        float x[NC][U], r[NC][U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = L::voxel(k0 + u);
            const long v = (long)f * VOX + (q < VOX ? q : VOX - 1);
            e[u] = gld(p.eps + (long)c * N + v);
            if constexpr (CORR) {
#pragma unroll
                for (int j = 0; j < NC; ++j) {
                    const long o = (long)(NORMAL ? j : c) * N + v;
                    x[j][u] = gld(p.corr + o);
                    r[j][u] = gld(rv + o);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = L::voxel(k0 + u);
            if (k0 + u >= K || q >= VOX) continue;
            double o = e[u] + sh_c;
            if constexpr (CORR) {
                if constexpr (NORMAL) {
                    const double xm[6] = {(double)x[0][u], (double)x[1][u], (double)x[2][u], 0.0, 0.0, 0.0};
                    const double rm[6] = {has_r ? (double)r[0][u] : 0.0, has_r ? (double)r[1][u] : 0.0,
                                          has_r ? (double)r[2][u] : 0.0, 0.0, 0.0, 0.0};
                    double tx[6], trr[6];
                    mandel_to_cart(xm, tx);
                    mandel_to_cart(rm, trr);
                    o += (c == 0) ? tx[0] : (c == 1) ? tx[1] : tx[2];
                    o += (c == 0) ? trr[0] : (c == 1) ? trr[1] : trr[2];
                } else {
                    o += (double)x[0][u] * kRsqrt2;
                    o += (has_r ? (double)r[0][u] : 0.0) * kRsqrt2;
                }
            }
            const long v = (long)f * VOX + q;
            const int m = q / NL, jl = q % NL;
            const double nw = bar + bufr[2 * ((m >> 1) * NLP + jl) + (m & 1)];
            const double d = nw - o;
            acc[0] += wgt * d * d;
            acc[1] += wgt * o * o;
            acc[2] += wgt * nw * nw;
            gst(p.eps + (long)c * N + v, nw);
            if (has_r) gst(p.p32 + (long)c * N + v, (float)d);   // the increment that starts a CG cycle (Cartesian)
        }
    }
}

template <int NP, bool BATCH>
__global__ __launch_bounds__(PlaneDims<NP>::T) void k_fwd_plane64(const LsParams p_arg, const LsParams* __restrict__ P, int flags) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    using PD = PlaneDims<NP>;
    constexpr int NL = PD::NL, NM = PD::NM, NLP = PD::NLP, VOX = PD::VOX;
    using cx = cplx<double>;
    if (p.ctl->done) return;
    cx* tw_l = reinterpret_cast<cx*>(smem_raw);
    cx* tw_m = tw_l + NL;
    cx* buf = tw_m + NM;
    cx* pl = buf + (NM / 2) * NLP;
    const Geom g = p.g;
    const int c = blockIdx.x / g.nf, f = blockIdx.x % g.nf;
    stage_twiddles<double>(tw_l, reinterpret_cast<const cx*>(p.tw_l64), NL);
    stage_twiddles<double>(tw_m, reinterpret_cast<const cx*>(p.tw_m64), NM);
    const bool add_shift = flags & LS_FLAG_SHIFT;
    const bool add_corr = (flags & LS_FLAG_CORR) && p.corr && p.ctl->corr_valid;
    double sh[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) sh[j] = add_shift ? p.ctl->shift[j] : 0.0;
    const double sh_c = add_shift ? p.ctl->shift[c] : 0.0;
    double* bufr = reinterpret_cast<double*>(buf);
    if (c < 3) {
        if (add_corr) plane64_tau<NP, true, true>(p, c, f, sh, sh_c, bufr);
        else plane64_tau<NP, true, false>(p, c, f, sh, sh_c, bufr);
    } else {
        if (add_corr) plane64_tau<NP, false, true>(p, c, f, sh, sh_c, bufr);
        else plane64_tau<NP, false, false>(p, c, f, sh, sh_c, bufr);
    }
    cx* spec = reinterpret_cast<cx*>(p.spec64) + (long)f * NM * PD::NLH * 6 + c * kPlaneTile;
    plane_fwd_fft<NP>(buf, pl, tw_l, tw_m, [spec](int idx, cx v) { gst(spec + tile_off(idx), v); }, (int)threadIdx.x);
}

template <int NP, bool BATCH>
__global__ __launch_bounds__(PlaneDims<NP>::T) void k_inv_plane64(const LsParams p_arg, const LsParams* __restrict__ P, int flags) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    using PD = PlaneDims<NP>;
    constexpr int NL = PD::NL, NM = PD::NM, NLP = PD::NLP, VOX = PD::VOX, SPT = PD::SPT;
    using cx = cplx<double>;
    if (p.ctl->done) return;
    cx* tw_l = reinterpret_cast<cx*>(smem_raw);
    cx* tw_m = tw_l + NL;
    cx* buf = tw_m + NM;
    cx* pl = buf + (NM / 2) * NLP;
    __shared__ double red[16 * 3];
    const Geom g = p.g;
    const int c = blockIdx.x / g.nf, f = blockIdx.x % g.nf;
    const cx* spec = reinterpret_cast<const cx*>(p.spec64) + (long)f * NM * PD::NLH * 6 + c * kPlaneTile;
    if constexpr (!BATCH) {
        cx sp[SPT];
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int idx = threadIdx.x + PlaneDims<NP>::T * k;
            sp[k] = gld(spec + tile_off(idx < PD::NS ? idx : 0));
        }
        stage_twiddles<double>(tw_l, reinterpret_cast<const cx*>(p.tw_l64), NL);
        stage_twiddles<double>(tw_m, reinterpret_cast<const cx*>(p.tw_m64), NM);
        plane_inv_fft<NP, double>(sp, buf, pl, tw_l, tw_m, (int)threadIdx.x);
    } else {
        // Batched launches read the twiddle pointers through the parameter
        // block, and a register copy of the spectrum ends up in scratch
        // there: the bins go straight to LDS instead (same placement).
        stage_twiddles<double>(tw_l, reinterpret_cast<const cx*>(p.tw_l64), NL);
        stage_twiddles<double>(tw_m, reinterpret_cast<const cx*>(p.tw_m64), NM);
        for (int idx = threadIdx.x; idx < PD::NS; idx += PlaneDims<NP>::T)
            pl[(idx % PD::NLH) * PD::NMP + idx / PD::NLH] = gld(spec + tile_off(idx));
        plane_inv_fft_staged<NP, double>(buf, pl, tw_l, tw_m, (int)threadIdx.x);
    }
    const bool add_shift = flags & LS_FLAG_SHIFT;
    const bool add_corr = (flags & LS_FLAG_CORR) && p.corr && p.ctl->corr_valid;
    const double bar = p.ctl->eps_bar[c];
    const double wgt = (c < 3) ? 1.0 : 2.0;   // off-diagonal entries appear twice in the full tensor
    const double sh_c = add_shift ? p.ctl->shift[c] : 0.0;
    const double* bufr = reinterpret_cast<const double*>(buf);
    double acc[3] = {0.0, 0.0, 0.0};
    if (c < 3) {
        if (add_corr) plane64_update<NP, true, true>(p, c, f, sh_c, bar, wgt, bufr, acc);
        else plane64_update<NP, true, false>(p, c, f, sh_c, bar, wgt, bufr, acc);
    } else {
        if (add_corr) plane64_update<NP, false, true>(p, c, f, sh_c, bar, wgt, bufr, acc);
        else plane64_update<NP, false, false>(p, c, f, sh_c, bar, wgt, bufr, acc);
    }
    block_sum<3>(acc, red);
    if (threadIdx.x == 0) {
        gst(p.partials + 3 * blockIdx.x, acc[0]);
        gst(p.partials + 3 * blockIdx.x + 1, acc[1]);
        gst(p.partials + 3 * blockIdx.x + 2, acc[2]);
    }
}

// ---------------------------------------------------------------------------
// Persistent CG cycle (one context on a cubic plane grid).  One launch whose
// blocks are all resident runs all CG steps of a refinement cycle.  Block x is (component,
// plane) as in the plane passes; its CG vectors r, w, p, s and the iterate x
// stay in registers for the whole cycle, so a step moves only the spectrum
// (out and back in once per phase) and the reduction partials, through
// device-coherent accesses.  A step's phases, between grid barriers:
//   F: CG update, tau_c, forward in-plane transforms, spectrum out;
//   G: blocks below the tile count run pass C on one tile each (F transform,
//      Green operator, inverse F transform);
//   I: inverse in-plane transforms, w_c = r_c - eps_tilde_c, partials out;
// then every block reduces all partials in the same order and takes the same
// decision.  The arithmetic and the reduction orders are the plane passes'
// (the folded reduction; k_ls_reduce's for the step enqueued last), so the
// iterates and the control block equal theirs bit for bit.  At the end the
// blocks write r and x back for the final pass or the next FP64 iteration.
// The grid barrier's words (p.gbar) are zeroed by the FP64 iteration's
// reduction that precedes every cycle.
// ---------------------------------------------------------------------------
// LDS of a cycle block: twiddles of the three axes, the plane passes' or
// pass C's buffers (one at a time), then the block's s and x (NP*NP floats
// each, 16-byte aligned).
template <int NP> constexpr size_t cycle_work_cplx() {
    constexpr int G = gamma_points<float, NP, 6>();
    constexpr size_t plane = (size_t)(NP / 2) * padded<NP>() + (size_t)(NP / 2 + 1) * padded<NP>();
    constexpr size_t gamma = (size_t)G * 6 * padded<NP>();
    return ((3 * NP + (plane > gamma ? plane : gamma)) + 1) / 2 * 2;
}
// The lane index as a value the compiler cannot move out of the step loop:
// addresses derived from it are recomputed in each phase instead of being
// held (and spilled) across the whole cycle.
__device__ __forceinline__ int lane_opaque() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

template <int NP> constexpr size_t cycle_lds_bytes() {
    return cycle_work_cplx<NP>() * sizeof(cplx<float>) + 2 * (size_t)NP * NP * sizeof(float);
}

template <int NP>
__global__ __launch_bounds__(PlaneDims<NP>::T) __attribute__((amdgpu_waves_per_eu(4))) void k_cg_cycle(const LsParams p, int J) {
    using PD = PlaneDims<NP>;
    constexpr int NF = NP;
    constexpr int NL = PD::NL, NM = PD::NM, NLP = PD::NLP, VOX = PD::VOX, V4 = PD::V4, Q4 = PD::Q4, SPT = PD::SPT,
                  PU = PD::PU;
    constexpr int G = gamma_points<float, NF, 6>();
    static_assert(gamma_threads<float, NF, 6>() == PlaneDims<NP>::T, "pass C runs on the plane blocks' lanes");
    using cx = cplx<float>;
    Ctl* ctl = p.ctl;
    // Every block reads the control block before its first barrier arrival;
    // block 0 writes it only after the first barrier.
    if ((ctl->done | ctl->cg_stop) != 0) return;
    const int cyc = cycle_len_now(ctl);
    const int steps = (cyc > 0 && cyc < J) ? cyc : J;
    CgIn st = cg_in(ctl, true, 0);   // step 0 counts on from slot 1, the cycle's start
    cx* tw_l = reinterpret_cast<cx*>(smem_raw);
    cx* tw_m = tw_l + NL;
    cx* tw_f = tw_m + NM;
    cx* buf = tw_f + NF;               // plane passes: NM/2 packed row pairs of NLP ...
    cx* pl = buf + (NM / 2) * NLP;     // ... and NLH columns of NMP; pass C: its lines from buf on
    float* s_lds = reinterpret_cast<float*>(reinterpret_cast<cx*>(smem_raw) + cycle_work_cplx<NP>());   // s, then x
    float* x_lds = s_lds + VOX;
    __shared__ double red[16 * 3];
    __shared__ double s_al, s_be;
    __shared__ int s_halt, s_flag;
    const Geom g = p.g;
    const int c = blockIdx.x / g.nf, f = blockIdx.x % g.nf;
    const long base = (long)c * g.N + (long)f * VOX;

    // r, w, p and the tau factor in registers; s and x in LDS (they are
    // touched once per step only)
    const vox4 zero4 = {0.f, 0.f, 0.f, 0.f};
    vox4 r[Q4], w[Q4], pv[Q4], tf[Q4];
#pragma unroll
    for (int k = 0; k < Q4; ++k) {
        const int q = threadIdx.x + PlaneDims<NP>::T * k;
        const int qq = q < V4 ? q : 0;
        r[k] = p.r0_cart ? r0_from_cart(p, c, (long)f * VOX, qq) : ld_vox4(p.r32 + base, qq);
        const vox4 dm = ld_vox4(p.dmu32 + (long)f * VOX, qq);
        const vox4 dl = (c == 0) ? ld_vox4(p.dlam32 + (long)f * VOX, qq) : zero4;
        tf[k] = tau_factor(c, dl, dm);
        w[k] = pv[k] = zero4;
        if (q < V4) {
            st_vox4_lds(s_lds, q, zero4);
            st_vox4_lds(x_lds, q, zero4);
        }
    }
    stage_twiddles<float>(tw_l, reinterpret_cast<const cx*>(p.tw_l32), NL);
    stage_twiddles<float>(tw_m, reinterpret_cast<const cx*>(p.tw_m32), NM);
    stage_twiddles<float>(tw_f, reinterpret_cast<const cx*>(p.tw_f32), NF);

    const long nxi = (long)g.nm * g.nlh;
    const CohBuf spec(p.spec32, (unsigned)((long)g.nf * nxi * 6 * sizeof(cx)));
    const CohBuf parts(p.partials, (unsigned)(3 * 6 * g.nf * sizeof(double)));
    const long plane_off = (long)f * NM * PD::NLH * 6 + c * kPlaneTile;
    const int ntiles = (int)(nxi / G);   // nxi is a multiple of G on plane grids
    const int nbp = 6 * g.nf;
    unsigned epoch = 0;
    float al = 0.f, be = 0.f;
    bool ok = true, updated = false;
    for (int j = 0; j < steps; ++j) {
        // F: update (from the previous step's decision), tau, forward in-plane transforms
        const bool upd = j >= 1, upd1 = j == 1;
        float* bufr = reinterpret_cast<float*>(buf);
        int tid = lane_opaque();
#pragma unroll
        for (int k = 0; k < Q4; ++k) {
            const int q = tid + PlaneDims<NP>::T * k;
            if (q >= V4) continue;
            vox4 sv = ld_vox4_lds(s_lds, q), xv = ld_vox4_lds(x_lds, q);
            plane_update_tau<NP>(q, upd, upd1, al, be, r[k], w[k], pv[k], sv, xv, tf[k], bufr);
            st_vox4_lds(s_lds, q, sv);
            st_vox4_lds(x_lds, q, xv);
        }
        updated = updated || upd;
        plane_fwd_fft<NP>(buf, pl, tw_l, tw_m, [&](int idx, cx v) { spec.stc(plane_off + tile_off(idx), v); }, tid);
        if (!(ok = grid_sync(p.gbar, ++epoch, &s_flag))) break;

        // G: pass C on tile blockIdx.x
        tid = lane_opaque();
        if ((int)blockIdx.x < ntiles) {
            const long xi0 = (long)blockIdx.x * G;
            const GammaLanes<float, NF, 6, true> ln(xi0, nxi, tid);
            const CohLine column{spec, ln.active ? xi0 * 6 + ln.l : 0, nxi * 6};
            cx v[NF / FftPlan<NF>::threads];
            line_gather<NF, float>(column, ln.t, ln.active, v);
            gamma_core<float, NF, 6, true>(v, ln, column, tw_f, buf, xi0, g, p.lam0, p.mu0, tid, PlaneDims<NP>::T);
        }
        if (!(ok = grid_sync(p.gbar, ++epoch, &s_flag))) break;

        // I: inverse in-plane transforms, w, partials
        tid = lane_opaque();
        cx sp[SPT];
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int idx = tid + PlaneDims<NP>::T * k;
            sp[k] = spec.ldc(plane_off + tile_off(idx < PD::NS ? idx : 0));
        }
        plane_inv_fft<NP>(sp, buf, pl, tw_l, tw_m, tid);
        double acc[3];
        plane_w_partials<NP>(reinterpret_cast<const float*>(buf), c, p.lam0, p.mu0, r, w, acc, tid);
        block_sum<3>(acc, red);
        if (threadIdx.x == 0)
#pragma unroll
            for (int k = 0; k < 3; ++k) parts.std_(3 * blockIdx.x + k, acc[k]);
        if (!(ok = grid_sync(p.gbar, ++epoch, &s_flag))) break;

        // the step's reduction and decision, the same in every block
        tid = lane_opaque();
        double tot[3];
        if (j == J - 1) {
            reduce_partials_ld<3>([&](long i) { return parts.ldd(i); }, nbp, 256, tot, red);
        } else {
            double part[3][PU];
#pragma unroll
            for (int u = 0; u < PU; ++u) {
                const int b = tid + PlaneDims<NP>::T * u;
#pragma unroll
                for (int k = 0; k < 3; ++k) part[k][u] = parts.ldd(3 * (b < nbp ? b : 0) + k);
            }
            plane_fold_sum<PU, PD::T>(part, nbp, tot);
            block_sum<3>(tot, red);
        }
        if (threadIdx.x == 0) {
            const CgOut o = cg_core(st, tot, j == 0);
            if (blockIdx.x == 0) cg_write(ctl, o, j == 0, true, j & 1, st.iters);
            s_al = o.al;
            s_be = o.be;
            s_halt = (o.done || o.stop) ? 1 : 0;
            st.inner_prev = o.inner;
            st.g_old = o.gam;
            st.a_old = o.al;
        }
        __syncthreads();
        al = (float)s_al;
        be = (float)s_be;
        if (s_halt) break;
    }
    if (!ok) {
        if (threadIdx.x == 0) ctl->cg_err = 1;
        return;
    }
    if (updated) {
#pragma unroll
        for (int k = 0; k < Q4; ++k) {
            const int q = threadIdx.x + PlaneDims<NP>::T * k;
            if (q >= V4) continue;
            st_vox4(p.r32 + base, q, r[k]);
            st_vox4(p.corr + base, q, ld_vox4_lds(x_lds, q));
        }
    }
}

// Reduction of pass E's block partials and the control-block update.  A kernel
// of its own: the kernel boundary makes every block's partials visible, with
// no per-block device-scope fence (which writes back the XCD's L2).  A
// last-block reduction inside pass E (release fence and ticket per block)
// measured 12% fewer FP32+refine events/s at 64^3 than this launch.
template <int MODE, bool BATCH>
__global__ __launch_bounds__(256) void k_ls_reduce(const LsParams p_arg, const LsParams* __restrict__ P, int nblocks, int flags) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    if (p.ctl->done || (MODE == LS_CG && (p.ctl->cg_stop || past_cycle(p.ctl, flags >> LS_STEP_SHIFT)))) return;
    if constexpr (MODE == LS_FULL) {
        // a persistent CG cycle may follow: its grid barrier starts from zero
        if (p.gbar)
            for (int i = threadIdx.x; i < kBarWords; i += blockDim.x) p.gbar[i] = 0u;
    }
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
            ctl->cycle_r0 = diff;
            if (ctl->iters >= 1 && ctl->inner > ctl->cycle_start) {
                // the CG cycle before this iteration: its length, and whether it
                // ran out of steps (no convergence, verification or breakdown stop)
                const int k = (ctl->iters == 1) ? 0 : 1;
                ctl->cyc_steps[k] = max(ctl->cyc_steps[k], ctl->inner - ctl->cycle_start);
                if (!ctl->cg_stop) ctl->cyc_exhausted[k] = 1;
            }
            ctl->cycle_start = ctl->inner;
            if (ctl->iters + ctl->inner < 32) ctl->diff_hist[ctl->iters + ctl->inner] = diff;
            ctl->iters += 1;
            ctl->corr_valid = 0;   // a pending correction was absorbed into eps
            ctl->cg_stop = 0;
            if (diff < ctl->tol || ctl->iters + ctl->inner >= ctl->max_iter) ctl->done = 1;
            ctl->cgs_inner[1] = ctl->inner;   // the count a cycle's first CG reduction starts from
        } else {
            const bool slots = flags & LS_FLAG_PLANE_SLOTS;
            cg_step_update(ctl, tot, (flags & LS_FLAG_CG_FIRST) != 0, slots, (flags & LS_FLAG_ODD) ? 1 : 0, true);
        }
    }
}


// ---------------------------------------------------------------------------
// Field mean -> warm-start shift (linear_elastic_solver.hpp:181-191); or fill
// with eps_bar for a cold start (:193).
// ---------------------------------------------------------------------------
template <int C, bool BATCH>
__global__ __launch_bounds__(256) void k_mean_shift(const LsParams p_arg, const LsParams* __restrict__ P) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
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

template <int C, bool BATCH>
__global__ __launch_bounds__(256) void k_mean_reduce(const LsParams p_arg, const LsParams* __restrict__ P, int nblocks) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    if (p.ctl->keep_shift) return;
    __shared__ double red[8 * 6];
    double tot[C];
    reduce_partials<C>(p.partials, nblocks, tot, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < C; ++c) p.ctl->shift[c] = p.ctl->eps_bar[c] - tot[c] / (double)p.g.N;
    }
}

template <int C, bool BATCH>
__global__ __launch_bounds__(256) void k_fill_bar(const LsParams p_arg, const LsParams* __restrict__ P) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    const Geom g = p.g;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < g.N; v += (long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < C; ++c) p.eps[c * g.N + v] = p.ctl->eps_bar[c];
}

// eps += shift (a warm start with no iteration still shifts, :181-191)
template <int C, bool BATCH>
__global__ __launch_bounds__(256) void k_apply_shift(const LsParams p_arg, const LsParams* __restrict__ P) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    const Geom g = p.g;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < g.N; v += (long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < C; ++c) p.eps[c * g.N + v] += p.ctl->shift[c];
}

// ---------------------------------------------------------------------------
// Final stress and macros (linear_elastic_solver.hpp:299-308, :583-592).
// ---------------------------------------------------------------------------
template <int C, bool BATCH>
__global__ __launch_bounds__(256) void k_final_stress(const LsParams p_arg, const LsParams* __restrict__ P) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    if (p.ctl->skip_final) return;
    constexpr int D = (C == 3) ? 2 : 3;
    __shared__ double red[8 * 12];
    const Geom g = p.g;
    double acc[2 * C];
#pragma unroll
    for (int c = 0; c < 2 * C; ++c) acc[c] = 0.0;
    // The solve ended inside an FP32 CG cycle: eps absorbs the iterate's
    // correction x and the residual r of the last step, which is the increment
    // the basic scheme would add at that iterate -- so the returned field
    // includes its last increment, as the reference's loop does (update, then
    // test).
    const bool absorb = p.corr && p.ctl->corr_valid;
    const bool has_epsp = p.epsp != nullptr;
    // absent optional fields read a valid stand-in, and every load of a voxel
    // is issued before its first store (the compiler keeps loads behind
    // stores to fields it cannot prove distinct)
    const float* src_cr = absorb ? p.corr : reinterpret_cast<const float*>(p.eps);
    const float* src_r = absorb ? p.r32 : reinterpret_cast<const float*>(p.eps);
    const double* src_ep = has_epsp ? p.epsp : p.eps;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < g.N; v += (long)gridDim.x * blockDim.x) {
        double e[C], ee[C], ep[C];
        float cr[C], rr[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            e[c] = p.eps[c * g.N + v];
            cr[c] = src_cr[c * g.N + v];
            rr[c] = src_r[c * g.N + v];
            ep[c] = src_ep[c * g.N + v];
        }
        const double mu_v = p.mu[v], lam_v = p.lam[v];
        double xc[C], rc[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            xc[c] = (double)cr[c];
            rc[c] = (double)rr[c];
        }
        if constexpr (C == 6) {
            if (absorb && p.mandel) {
                double t[6];
                mandel_to_cart(xc, t);
#pragma unroll
                for (int c = 0; c < 6; ++c) xc[c] = t[c];
                mandel_to_cart(rc, t);
#pragma unroll
                for (int c = 0; c < 6; ++c) rc[c] = t[c];
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            if (absorb) {
                e[c] += xc[c];
                e[c] += rc[c];
                p.eps[c * g.N + v] = e[c];
            }
            ee[c] = has_epsp ? e[c] - ep[c] : e[c];
        }
        double tre = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) tre += ee[c];
        const double tm = 2.0 * mu_v, lt = lam_v * tre;
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

template <int C, bool BATCH>
__global__ __launch_bounds__(256) void k_final_reduce(const LsParams p_arg, const LsParams* __restrict__ P, int nblocks) {
    const LsParams& p = param_sel<BATCH>(p_arg, P);
    if (p.ctl->skip_final) return;
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

template <typename SR, int NN, int C, int MODE, bool BATCH, int HALF>
static void launch_fwd_last_v(const LsLaunch& p, int flags, hipStream_t s) {
    constexpr int pairs = fwd_pairs_v<SR, NN, C, HALF>();
    const size_t sm = (size_t)(NN + C * pairs * padded<NN>()) * sizeof(cplx<SR>);
    hipLaunchKernelGGL((k_fwd_last<SR, NN, C, MODE, BATCH, HALF>), dim3(blocks_for(p.g.nlines(), 2 * pairs), p.B),
                       dim3(256), sm, s, p.h, p.P, flags);
}

// Variant 1 exists for 3D lines of 32 points or more (the tuned grids).
template <int C, int NN> constexpr bool tuned_len() { return C == 6 && NN >= 32; }

template <typename SR, int C, int MODE, bool BATCH>
static void launch_fwd_last(const LsLaunch& p, int flags, hipStream_t s) {
    const int nl = p.g.nl;
    const int var = p.tune.fwd[MODE == LS_FULL ? 0 : 1];
#define MGK_CASE(NN)                                                                                          \
    if (nl == NN) {                                                                                           \
        if constexpr (tuned_len<C, NN>()) {                                                                   \
            if (var) { launch_fwd_last_v<SR, NN, C, MODE, BATCH, 1>(p, flags, s); return; }                  \
        }                                                                                                     \
        launch_fwd_last_v<SR, NN, C, MODE, BATCH, 0>(p, flags, s);                                            \
        return;                                                                                               \
    }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the last axis");
}

template <typename SR, int C, int MODE, bool BATCH>
static void launch_inv_last(const LsLaunch& p, int flags, hipStream_t s) {
    const int nl = p.g.nl;
#define MGK_CASE(NN)                                                                                  \
    if (nl == NN) {                                                                                   \
        constexpr int pairs = inv_pairs<SR, NN, C>();                                                \
        const size_t sm = (size_t)(NN + C * pairs * padded<NN>()) * sizeof(cplx<SR>);                \
        const int nb = blocks_for(p.g.nlines(), 2 * pairs);                                           \
        hipLaunchKernelGGL((k_inv_last<SR, NN, C, MODE, BATCH>), dim3(nb, p.B), dim3(256), sm, s, p.h, p.P, flags);     \
        hipLaunchKernelGGL((k_ls_reduce<MODE, BATCH>), dim3(1, p.B), dim3(256), 0, s, p.h, p.P, nb, flags);              \
        return;                                                                                       \
    }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the last axis");
}

template <typename SR, int NN, int DIR, bool BATCH, int HALF>
static void launch_mid_v(const LsLaunch& p, hipStream_t s, int mode, int spec, int step) {
    const long total = (long)p.g.nf * p.g.nlh * p.g.C;
    constexpr int L = mid_cols_v<SR, NN, HALF>();
    const size_t sm = (size_t)(NN + L * padded<NN>()) * sizeof(cplx<SR>);
    hipLaunchKernelGGL((k_mid<SR, NN, DIR, BATCH, HALF>), dim3(blocks_for(total, L), p.B), dim3(L * FftPlan<NN>::threads),
                       sm, s, p.h, p.P, mode, spec, step);
}

template <typename SR, int DIR, bool BATCH>
static void launch_mid(const LsLaunch& p, hipStream_t s, int mode, int spec = 0, int step = 0) {
    const int nm = p.g.nm;
    const int var = p.tune.mid[mode == LS_FULL ? 0 : 1];
#define MGK_CASE(NN)                                                                                       \
    if (nm == NN) {                                                                                        \
        if constexpr (tuned_len<6, NN>()) {                                                                \
            if (var) { launch_mid_v<SR, NN, DIR, BATCH, 1>(p, s, mode, spec, step); return; }             \
        }                                                                                                  \
        launch_mid_v<SR, NN, DIR, BATCH, 0>(p, s, mode, spec, step);                                       \
        return;                                                                                            \
    }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the middle axis");
}

template <typename SR, int C, bool BATCH, bool PLANE = false>
static void launch_first_gamma(const LsLaunch& p, hipStream_t s, int mode, int spec = 0, int step = 0) {
    const int nf = p.g.nf;
    const long nxi = (long)p.g.nm * p.g.nlh;
#define MGK_CASE(NN)                                                                                 \
    if (nf == NN) {                                                                                  \
        if constexpr (PLANE && std::is_same<SR, float>::value && C == 6 && NN == 128) {              \
            constexpr int LW = plane_gamma_lanes<NN>();                                              \
            constexpr int G = gamma_points<float, NN, 6, LW>();                                      \
            const size_t sm = (size_t)(NN + G * 6 * padded<NN>()) * sizeof(cplx<float>);             \
            hipLaunchKernelGGL((k_first_gamma_wide<NN, BATCH>), dim3(blocks_for(nxi, G), p.B),         \
                               dim3(gamma_threads<float, NN, 6, LW>()), sm, s, p.h, p.P, spec, step);   \
            return;                                                                                  \
        }                                                                                            \
        constexpr int G = gamma_points<SR, NN, C>();                                                 \
        const size_t sm = (size_t)(NN + G * C * padded<NN>()) * sizeof(cplx<SR>);                    \
        hipLaunchKernelGGL((k_first_gamma<SR, NN, C, BATCH, PLANE>), dim3(blocks_for(nxi, G), p.B),              \
                           dim3(gamma_threads<SR, NN, C>()), sm, s, p.h, p.P, mode, spec, step);           \
        return;                                                                                      \
    }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the first axis");
}

// Doubles of block partials: the largest grid of any reducing kernel times
// its partial width (pass E runs at least two lines per block, the final pass
// at most 1024 blocks), rounded to whole 128-byte lines.
static long partial_rows(const Geom& g) {
    long lines = g.nlines();
    long nb = (lines + 1) / 2;
    return (std::max(nb * 3, 1024L * 12) + 15) / 16 * 16;
}

// The partials buffer ends with the persistent CG cycle's grid barrier words.
int ls_partial_slots(const Geom& g) { return (int)(partial_rows(g) + kBarWords / 2); }
long ls_gbar_offset(const Geom& g) { return partial_rows(g); }

template <bool BATCH> static void launch_fwd_plane64(const LsLaunch& p, int flags, hipStream_t s);
template <bool BATCH> static void launch_first_gamma64_tiled(const LsLaunch& p, hipStream_t s);
template <bool BATCH> static void launch_inv_plane64(const LsLaunch& p, int flags, hipStream_t s, bool reduce);

template <int C, bool BATCH> static void iteration_full(const LsLaunch& p, int flags, hipStream_t s) {
    if constexpr (C == 6) {
        if (p.plane64) {
            launch_fwd_plane64<BATCH>(p, flags, s);
            launch_first_gamma64_tiled<BATCH>(p, s);
            launch_inv_plane64<BATCH>(p, flags, s, true);
            return;
        }
    }
    launch_fwd_last<double, C, LS_FULL, BATCH>(p, flags, s);
    if (C == 6 && p.g.nm > 1) launch_mid<double, -1, BATCH>(p, s, LS_FULL);
    launch_first_gamma<double, C, BATCH>(p, s, LS_FULL);
    if (C == 6 && p.g.nm > 1) launch_mid<double, +1, BATCH>(p, s, LS_FULL);
    launch_inv_last<double, C, LS_FULL, BATCH>(p, flags, s);
}

#define MGK_PLANE_CASES(X) X(8) X(16) X(32) X(48) X(64) X(128)

bool ls_plane_path(const Geom& g) {
    const char* e = std::getenv("MGK_LS_PLANE");   // read per launch set-up: a test can compare both paths
    return !(e && e[0] == '0') && ls_plane_supported(g);
}

static bool plane_gamma_tiles(int nf) {
#define MGK_CASE(NN) if (nf == NN) return gamma_points<float, NN, 6, plane_gamma_lanes<NN>()>() % kPlaneTile == 0;
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    return false;
}

bool ls_plane_supported(const Geom& g) {
    if (g.C != 6 || g.nm <= 1 || g.nm != g.nl) return false;
    if (((long)g.nm * g.nlh) % kPlaneTile != 0 || !plane_gamma_tiles(g.nf)) return false;
#define MGK_CASE(NN) if (g.nm == NN) return true;
    MGK_PLANE_CASES(MGK_CASE)
#undef MGK_CASE
    return false;
}

// Kernels whose dynamic LDS exceeds the default 64 KiB limit are allowed
// their size once (128-point planes: 132 KiB).
template <class K> static void allow_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024) MGK_HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

template <int NN, bool BATCH, int PU> static void launch_fwd_plane_pu(const LsLaunch& p, int flags, hipStream_t s) {
    static const bool lds_ok = (allow_lds(k_fwd_plane<NN, BATCH, PU>, plane_lds_bytes<NN>()), true);
    (void)lds_ok;
    hipLaunchKernelGGL((k_fwd_plane<NN, BATCH, PU>), dim3(6 * p.g.nf, p.B), dim3(PlaneDims<NN>::T), plane_lds_bytes<NN>(),
                       s, p.h, p.P, flags);
}

template <bool BATCH> static void launch_fwd_plane(const LsLaunch& p, int flags, hipStream_t s) {
#define MGK_CASE(NN)                                                                                  \
    if (p.g.nm == NN) {                                                                               \
        if (6 * p.g.nf <= PlaneDims<NN>::T) launch_fwd_plane_pu<NN, BATCH, 1>(p, flags, s);          \
        else launch_fwd_plane_pu<NN, BATCH, PlaneDims<NN>::PU>(p, flags, s);                          \
        return;                                                                                       \
    }
    MGK_PLANE_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("plane length not supported");
}

template <bool BATCH> static void launch_inv_plane(const LsLaunch& p, int flags, hipStream_t s, bool reduce = true) {
    const int nb = 6 * p.g.nf;
#define MGK_CASE(NN)                                                                                  \
    if (p.g.nm == NN) {                                                                               \
        static const bool lds_ok = (allow_lds(k_inv_plane<NN, BATCH>, plane_lds_bytes<NN>()), true);  \
        (void)lds_ok;                                                                                 \
        hipLaunchKernelGGL((k_inv_plane<NN, BATCH>), dim3(nb, p.B), dim3(PlaneDims<NN>::T),              \
                           plane_lds_bytes<NN>(), s, p.h, p.P, flags);                               \
        if (reduce)                                                                                   \
            hipLaunchKernelGGL((k_ls_reduce<LS_CG, BATCH>), dim3(1, p.B), dim3(256), 0, s, p.h, p.P, nb, flags); \
        return;                                                                                       \
    }
    MGK_PLANE_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("plane length not supported");
}

// Persistent CG cycles: cubic plane grids of these lengths, when the
// device holds all 6 n blocks at once (decided once per length).  Measured
// per CG step on MI355X: 28.6 us at 64^3 against 31.4 us for the three
// plane-pass launches (FP32+refine 2460 against 2298 events/s); at 32^3 and
// 48^3 the two are within 2% (the launched passes slightly ahead), so only
// 64^3 takes the persistent cycle.
#define MGK_CYCLE_CASES(X) X(64)

template <int NP> static bool cycle_fits() {
    static std::atomic<int> state{0};   // 0 unknown, 1 fits, 2 does not
    if (state.load() == 0) {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t pr;
        bool ok = hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess &&
                  hipFuncSetAttribute((const void*)k_cg_cycle<NP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)cycle_lds_bytes<NP>()) == hipSuccess &&
                  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_cg_cycle<NP>, PlaneDims<NP>::T,
                                                               cycle_lds_bytes<NP>()) == hipSuccess;
        ok = ok && (long)per_cu * pr.multiProcessorCount >= 6L * NP;
        state.store(ok ? 1 : 2);
    }
    return state.load() == 1;
}

bool ls_cycle_supported(const Geom& g) {
    const char* e = std::getenv("MGK_LS_CYCLE");   // MGK_LS_CYCLE=0: CG steps as separate plane passes
    if ((e && e[0] == '0') || !ls_plane_path(g) || g.nf != g.nm) return false;
#define MGK_CASE(NN) if (g.nm == NN) return cycle_fits<NN>();
    MGK_CYCLE_CASES(MGK_CASE)
#undef MGK_CASE
    return false;
}

// An ordinary launch whose blocks the device holds at once (cycle_fits); the
// caller keeps a second cycle from running beside it (one cycle in flight per
// process, see g_cycle_mu).
void ls_enqueue_cycle(const LsLaunch& p, int J, hipStream_t s) {
    if (p.batch || !p.h.gbar) throw ArgError("persistent CG cycles run on a single context");
#define MGK_CASE(NN)                                                                                           \
    if (p.g.nm == NN) {                                                                                        \
        hipLaunchKernelGGL((k_cg_cycle<NN>), dim3(6 * NN), dim3(PlaneDims<NN>::T), cycle_lds_bytes<NN>(), s, p.h, J); \
        MGK_HIP_CHECK(hipGetLastError());                                                                      \
        return;                                                                                                \
    }
    MGK_CYCLE_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("persistent CG cycle: grid not supported");
}


#define MGK_PLANE64_CASES(X) X(16) X(32) X(48) X(64)

// FP64 pass C on the tiled spectrum needs whole tiles per block.
static bool plane64_gamma_tiles(int nf) {
#define MGK_CASE(NN) if (nf == NN) return gamma_points<double, NN, 6, 512, true>() % kPlaneTile == 0;
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    return false;
}

bool ls_plane64_path(const Geom& g) {
    const char* e = std::getenv("MGK_LS_PLANE64");
    if ((e && e[0] == '0') || !ls_plane_path(g) || !plane64_gamma_tiles(g.nf)) return false;
#define MGK_CASE(NN) if (g.nm == NN) return true;
    MGK_PLANE64_CASES(MGK_CASE)
#undef MGK_CASE
    return false;
}

// FP64 pass C on the tiled spectrum: 512 lanes, whole tiles per block.
template <bool BATCH> static void launch_first_gamma64_tiled(const LsLaunch& p, hipStream_t s) {
    const long nxi = (long)p.g.nm * p.g.nlh;
#define MGK_CASE(NN)                                                                                          \
    if (p.g.nf == NN) {                                                                                       \
        constexpr int G = gamma_points<double, NN, 6, 512, true>();                                           \
        if constexpr (G % kPlaneTile == 0) {                                                                  \
            const size_t sm = (size_t)(NN + G * 6 * padded<NN>()) * sizeof(cplx<double>);                     \
            hipLaunchKernelGGL((k_first_gamma<double, NN, 6, BATCH, true, 512>), dim3(blocks_for(nxi, G), p.B), \
                               dim3(gamma_threads<double, NN, 6, 512, true>()), sm, s, p.h, p.P, LS_FULL, 0, 0);  \
            return;                                                                                           \
        }                                                                                                     \
    }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FP64 plane pass C: first-axis length without whole tiles");
}

template <bool BATCH> static void launch_fwd_plane64(const LsLaunch& p, int flags, hipStream_t s) {
#define MGK_CASE(NN)                                                                                  \
    if (p.g.nm == NN) {                                                                               \
        static const bool lds_ok = (allow_lds(k_fwd_plane64<NN, BATCH>, plane64_lds_bytes<NN>()), true); \
        (void)lds_ok;                                                                                 \
        hipLaunchKernelGGL((k_fwd_plane64<NN, BATCH>), dim3(6 * p.g.nf, p.B), dim3(PlaneDims<NN>::T),     \
                           plane64_lds_bytes<NN>(), s, p.h, p.P, flags);                             \
        return;                                                                                       \
    }
    MGK_PLANE64_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FP64 plane length not supported");
}

template <bool BATCH> static void launch_inv_plane64(const LsLaunch& p, int flags, hipStream_t s, bool reduce) {
    const int nb = 6 * p.g.nf;
#define MGK_CASE(NN)                                                                                  \
    if (p.g.nm == NN) {                                                                               \
        static const bool lds_ok = (allow_lds(k_inv_plane64<NN, BATCH>, plane64_lds_bytes<NN>()), true); \
        (void)lds_ok;                                                                                 \
        hipLaunchKernelGGL((k_inv_plane64<NN, BATCH>), dim3(nb, p.B), dim3(PlaneDims<NN>::T),            \
                           plane64_lds_bytes<NN>(), s, p.h, p.P, flags);                             \
        if (reduce)                                                                                   \
            hipLaunchKernelGGL((k_ls_reduce<LS_FULL, BATCH>), dim3(1, p.B), dim3(256), 0, s, p.h, p.P, nb, flags); \
        return;                                                                                       \
    }
    MGK_PLANE64_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FP64 plane length not supported");
}

template <int C, bool BATCH> static void iteration_delta(const LsLaunch& p, int flags, hipStream_t s) {
    const int spec = (flags & LS_FLAG_SPEC) ? 1 : 0;
    const int step = flags >> LS_STEP_SHIFT;
    if constexpr (C == 6) {
        if (p.plane) {
            // the next step's forward plane pass folds in this step's
            // reduction; the cycle's last step keeps a reduction kernel
            launch_fwd_plane<BATCH>(p, flags, s);
            launch_first_gamma<float, 6, BATCH, true>(p, s, LS_CG, spec, step);
            launch_inv_plane<BATCH>(p, flags | LS_FLAG_PLANE_SLOTS, s, (flags & LS_FLAG_LAST) != 0);
            return;
        }
    }
    launch_fwd_last<float, C, LS_CG, BATCH>(p, flags, s);
    if (C == 6 && p.g.nm > 1) launch_mid<float, -1, BATCH>(p, s, LS_CG, spec, step);
    launch_first_gamma<float, C, BATCH>(p, s, LS_CG, spec, step);
    if (C == 6 && p.g.nm > 1) launch_mid<float, +1, BATCH>(p, s, LS_CG, spec, step);
    launch_inv_last<float, C, LS_CG, BATCH>(p, flags, s);
}

static int reduce_grid(long n) {
    long nb = (n + 255) / 256;
    return (int)std::min<long>(std::max<long>(nb, 1), 1024);
}

template <bool BATCH> static void enqueue_iteration_t(const LsLaunch& p, int mode, int flags, hipStream_t s) {
    if (p.g.C == 3) {
        if (mode == LS_FULL) iteration_full<3, BATCH>(p, flags, s); else iteration_delta<3, BATCH>(p, flags, s);
    } else {
        if (mode == LS_FULL) iteration_full<6, BATCH>(p, flags, s); else iteration_delta<6, BATCH>(p, flags, s);
    }
}

void ls_enqueue_iteration(const LsLaunch& p, int mode, int flags, hipStream_t s) {
    if (p.batch) enqueue_iteration_t<true>(p, mode, flags, s); else enqueue_iteration_t<false>(p, mode, flags, s);
    MGK_HIP_CHECK(hipGetLastError());
}

template <typename SR, int C, int MODE, bool BATCH>
static void launch_inv_last_only(const LsLaunch& p, int flags, hipStream_t s) {
    const int nl = p.g.nl;
#define MGK_CASE(NN)                                                                              \
    if (nl == NN) {                                                                               \
        constexpr int pairs = inv_pairs<SR, NN, C>();                                            \
        const size_t sm = (size_t)(NN + C * pairs * padded<NN>()) * sizeof(cplx<SR>);            \
        hipLaunchKernelGGL((k_inv_last<SR, NN, C, MODE, BATCH>), dim3(blocks_for(p.g.nlines(), 2 * pairs), p.B), \
                           dim3(256), sm, s, p.h, p.P, flags);                                           \
        return;                                                                                   \
    }
    MGK_LEN_CASES(MGK_CASE)
#undef MGK_CASE
    throw UnsupportedError("FFT length not supported on the last axis");
}

template <int C, bool BATCH> static void enqueue_pass_c(const LsLaunch& p, int pass, int mode, hipStream_t s) {
    const bool full = (mode == LS_FULL);
    switch (pass) {
        case 0:
            if (full) launch_fwd_last<double, C, LS_FULL, BATCH>(p, 0, s); else launch_fwd_last<float, C, LS_CG, BATCH>(p, 0, s);
            break;
        case 1:
            if (p.g.nm == 1) throw ArgError("pass not present for this grid");
            if (full) launch_mid<double, -1, BATCH>(p, s, mode); else launch_mid<float, -1, BATCH>(p, s, mode); break;
        case 2: if (full) launch_first_gamma<double, C, BATCH>(p, s, mode); else launch_first_gamma<float, C, BATCH>(p, s, mode); break;
        case 3:
            if (p.g.nm == 1) throw ArgError("pass not present for this grid");
            if (full) launch_mid<double, +1, BATCH>(p, s, mode); else launch_mid<float, +1, BATCH>(p, s, mode); break;
        case 4:
            if (full) launch_inv_last_only<double, C, LS_FULL, BATCH>(p, 0, s);
            else launch_inv_last_only<float, C, LS_CG, BATCH>(p, 0, s);
            break;
        case 5:
            if (full) throw ArgError("the CG update pass runs in FP32 only");
            launch_fwd_last<float, C, LS_CG, BATCH>(p, LS_FLAG_CG_UPDATE, s);
            break;
        case 9:
        case 10:
        case 11:
            if constexpr (C == 6) {
                if (!full || !ls_plane64_path(p.g)) throw ArgError("FP64 plane passes run in FP64 on supported 3D grids");
                if (pass == 9) launch_fwd_plane64<BATCH>(p, 0, s);
                else if (pass == 10) launch_first_gamma64_tiled<BATCH>(p, s);
                else launch_inv_plane64<BATCH>(p, 0, s, false);
            } else {
                throw ArgError("plane passes are 3D only");
            }
            break;
        case 6:
        case 7:
        case 8:
            if constexpr (C == 6) {
                if (full || !ls_plane_supported(p.g)) throw ArgError("plane passes run in FP32 on supported 3D grids");
                if (pass == 6) launch_fwd_plane<BATCH>(p, LS_FLAG_CG_UPDATE | LS_FLAG_PROBE, s);
                else if (pass == 7) launch_first_gamma<float, 6, BATCH, true>(p, s, LS_CG);
                else launch_inv_plane<BATCH>(p, 0, s, false);
            } else {
                throw ArgError("plane passes are 3D only");
            }
            break;
        default: throw ArgError("unknown pass");
    }
}

void ls_enqueue_pass(const LsLaunch& p, int pass, int mode, hipStream_t s) {
    if (p.g.C == 3) {
        if (pass == 1 || pass == 3) throw ArgError("pass not present in 2D");
        if (p.batch) enqueue_pass_c<3, true>(p, pass, mode, s); else enqueue_pass_c<3, false>(p, pass, mode, s);
    } else {
        if (p.batch) enqueue_pass_c<6, true>(p, pass, mode, s); else enqueue_pass_c<6, false>(p, pass, mode, s);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

template <bool BATCH> static void enqueue_start_t(const LsLaunch& p, int start, hipStream_t s) {
    const int nb = reduce_grid(p.g.N);
    if (start == LS_START_KEEP) return;
    if (p.g.C == 3) {
        if (start == LS_START_MEAN) {
            hipLaunchKernelGGL((k_mean_shift<3, BATCH>), dim3(nb, p.B), dim3(256), 0, s, p.h, p.P);
            hipLaunchKernelGGL((k_mean_reduce<3, BATCH>), dim3(1, p.B), dim3(256), 0, s, p.h, p.P, nb);
        }
        else if (start == LS_START_FILL) hipLaunchKernelGGL((k_fill_bar<3, BATCH>), dim3(nb, p.B), dim3(256), 0, s, p.h, p.P);
        else hipLaunchKernelGGL((k_apply_shift<3, BATCH>), dim3(nb, p.B), dim3(256), 0, s, p.h, p.P);
    } else {
        if (start == LS_START_MEAN) {
            hipLaunchKernelGGL((k_mean_shift<6, BATCH>), dim3(nb, p.B), dim3(256), 0, s, p.h, p.P);
            hipLaunchKernelGGL((k_mean_reduce<6, BATCH>), dim3(1, p.B), dim3(256), 0, s, p.h, p.P, nb);
        }
        else if (start == LS_START_FILL) hipLaunchKernelGGL((k_fill_bar<6, BATCH>), dim3(nb, p.B), dim3(256), 0, s, p.h, p.P);
        else hipLaunchKernelGGL((k_apply_shift<6, BATCH>), dim3(nb, p.B), dim3(256), 0, s, p.h, p.P);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

template <bool BATCH> static void enqueue_final_t(const LsLaunch& p, hipStream_t s) {
    const int nb = reduce_grid(p.g.N);
    if (p.g.C == 3) {
        hipLaunchKernelGGL((k_final_stress<3, BATCH>), dim3(nb, p.B), dim3(256), 0, s, p.h, p.P);
        hipLaunchKernelGGL((k_final_reduce<3, BATCH>), dim3(1, p.B), dim3(256), 0, s, p.h, p.P, nb);
    } else {
        hipLaunchKernelGGL((k_final_stress<6, BATCH>), dim3(nb, p.B), dim3(256), 0, s, p.h, p.P);
        hipLaunchKernelGGL((k_final_reduce<6, BATCH>), dim3(1, p.B), dim3(256), 0, s, p.h, p.P, nb);
    }
    MGK_HIP_CHECK(hipGetLastError());
}

void ls_enqueue_start(const LsLaunch& p, int start, hipStream_t s) {
    if (p.batch) enqueue_start_t<true>(p, start, s); else enqueue_start_t<false>(p, start, s);
}

void ls_enqueue_final(const LsLaunch& p, hipStream_t s) {
    if (p.batch) enqueue_final_t<true>(p, s); else enqueue_final_t<false>(p, s);
}

}  // namespace mgk
