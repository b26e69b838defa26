// fft_lds.hpp -- LDS-resident mixed-radix (2/3/4/5/8) Stockham FFT for gfx950.
//
// A workgroup holds a batch of complex lines of one compile-time length N in
// LDS.  Each line is served by T = plan.threads lanes (lds_fft: the line
// index runs fastest across lanes, the column passes alike); in every
// stage a lane loads its V = N/T values into registers, the workgroup meets at
// a barrier, and the lane writes its butterfly outputs back (single buffer,
// read-all-then-write-all).  Twiddles come from a per-length table W_N^k that
// the caller has staged in LDS (exact long-double values prepared on the host).
//
// Conventions: forward = exp(-2 pi i jk/N), inverse = exp(+2 pi i jk/N), both
// unnormalised; the Lippmann-Schwinger pipeline applies 1/N once (the
// reference's ortho 1/sqrt(N) twice, pocketfft_hdronly.h:3383-3394).
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace mgk {

template <typename R> struct cplx { R x, y; };

template <typename R> __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R> __device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> __device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R> __device__ __forceinline__ cplx<R> cscale(cplx<R> a, R s) { return {a.x * s, a.y * s}; }
// a * (-i) for forward (DIR = -1), a * (+i) for inverse (DIR = +1)
template <int DIR, typename R> __device__ __forceinline__ cplx<R> mul_i(cplx<R> a) {
    return DIR < 0 ? cplx<R>{a.y, -a.x} : cplx<R>{-a.y, a.x};
}

// ---------------------------------------------------------------------------
// Compile-time plan: radices in stage order, threads per line.
// ---------------------------------------------------------------------------
template <int N> struct FftPlan;

#define MGK_PLAN(NN, TT, ...)                                                   \
    template <> struct FftPlan<NN> {                                            \
        static constexpr int n = NN;                                            \
        static constexpr int threads = TT;                                      \
        static constexpr int radix[] = {__VA_ARGS__};                           \
        static constexpr int stages = sizeof(radix) / sizeof(int);              \
    };

// threads T must divide N / radix for every stage
MGK_PLAN(2, 1, 2)
MGK_PLAN(4, 1, 4)
MGK_PLAN(6, 1, 2, 3)
MGK_PLAN(8, 1, 8)
MGK_PLAN(10, 1, 2, 5)
MGK_PLAN(12, 1, 4, 3)
MGK_PLAN(16, 2, 4, 4)
MGK_PLAN(20, 1, 4, 5)
MGK_PLAN(24, 1, 8, 3)
MGK_PLAN(32, 4, 8, 4)
MGK_PLAN(40, 1, 8, 5)
MGK_PLAN(48, 4, 4, 4, 3)
MGK_PLAN(64, 8, 8, 8)
MGK_PLAN(80, 4, 4, 4, 5)
MGK_PLAN(96, 4, 8, 4, 3)
MGK_PLAN(128, 16, 8, 8, 2)
MGK_PLAN(160, 4, 8, 4, 5)
MGK_PLAN(192, 8, 8, 8, 3)
MGK_PLAN(256, 32, 8, 8, 4)
#undef MGK_PLAN

// ---------------------------------------------------------------------------
// Small DFT butterflies on registers.
// ---------------------------------------------------------------------------
template <int R, int DIR, typename T> struct Bfly;

template <int DIR, typename T> struct Bfly<2, DIR, T> {
    static __device__ __forceinline__ void run(cplx<T>* v) {
        cplx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR, typename T> struct Bfly<4, DIR, T> {
    static __device__ __forceinline__ void run(cplx<T>* v) {
        cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        cplx<T> t2 = cadd(v[1], v[3]), t3 = mul_i<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int DIR, typename T> struct Bfly<8, DIR, T> {
    static __device__ __forceinline__ void run(cplx<T>* v) {
        cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Bfly<4, DIR, T>::run(e);
        Bfly<4, DIR, T>::run(o);
        const T h = (T)0.70710678118654752440084436210484904;
        // W8^1 = (1 -/+ i)/sqrt2, W8^2 = -/+ i, W8^3 = (-1 -/+ i)/sqrt2
        cplx<T> w1o = DIR < 0 ? cplx<T>{(o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h}
                              : cplx<T>{(o[1].x - o[1].y) * h, (o[1].y + o[1].x) * h};
        cplx<T> w2o = mul_i<DIR>(o[2]);
        cplx<T> w3o = DIR < 0 ? cplx<T>{(-o[3].x + o[3].y) * h, (-o[3].y - o[3].x) * h}
                              : cplx<T>{(-o[3].x - o[3].y) * h, (-o[3].y + o[3].x) * h};
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], w1o);
        v[5] = csub(e[1], w1o);
        v[2] = cadd(e[2], w2o);
        v[6] = csub(e[2], w2o);
        v[3] = cadd(e[3], w3o);
        v[7] = csub(e[3], w3o);
    }
};

template <int DIR, typename T> struct Bfly<3, DIR, T> {
    static __device__ __forceinline__ void run(cplx<T>* v) {
        const T s3 = (T)0.86602540378443864676372317075293618;   // sqrt(3)/2
        cplx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        cplx<T> m = {v[0].x - (T)0.5 * s.x, v[0].y - (T)0.5 * s.y};
        cplx<T> jd = mul_i<DIR>(cscale(d, s3));   // -/+ i (sqrt3/2) d
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, jd);
        v[2] = csub(m, jd);
    }
};

template <int DIR, typename T> struct Bfly<5, DIR, T> {
    static __device__ __forceinline__ void run(cplx<T>* v) {
        const T c1 = (T)0.30901699437494742410229341718281906;    // cos(2pi/5)
        const T c2 = (T)-0.80901699437494742410229341718281906;   // cos(4pi/5)
        const T s1 = (T)0.95105651629515357211643933337938214;    // sin(2pi/5)
        const T s2 = (T)0.58778525229247312916870595463907277;    // sin(4pi/5)
        cplx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        cplx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        cplx<T> x0 = v[0];
        cplx<T> p1 = {x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
        cplx<T> p2 = {x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
        cplx<T> q1 = mul_i<DIR>(cplx<T>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
        cplx<T> q2 = mul_i<DIR>(cplx<T>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
        v[0] = cadd(x0, cadd(a1, a2));
        v[1] = cadd(p1, q1);
        v[4] = csub(p1, q1);
        v[2] = cadd(p2, q2);
        v[3] = csub(p2, q2);
    }
};

// ---------------------------------------------------------------------------
// Line accessors.  A stage reads element j of its line through in(j) and
// writes element k through out(k, v): LDS lines, or global columns with a
// stride (the first and last stages of a register-resident transform).
// ---------------------------------------------------------------------------
template <typename R> struct LineRef {
    cplx<R>* p;
    long es;
    __device__ __forceinline__ cplx<R> operator()(int j) const { return p[j * es]; }
    __device__ __forceinline__ void operator()(int j, cplx<R> v) const { p[j * es] = v; }
};
// A line in device (global) memory: global, not flat, accesses also where the
// base pointer comes from a parameter block in memory.
template <typename R> struct GLineRef {
    cplx<R>* p;
    long es;
    __device__ __forceinline__ cplx<R> operator()(int j) const { return gld(p + j * es); }
    __device__ __forceinline__ void operator()(int j, cplx<R> v) const { gst(p + j * es, v); }
};

// ---------------------------------------------------------------------------
// One Stockham stage for one line: input element j + r*N/R, output
// (j/Ns)*Ns*R + j%Ns + r*Ns, twiddle W_N^{r*(j%Ns)*N/(Ns*R)}.
// ---------------------------------------------------------------------------
template <int N, int R, int NS, int DIR, int T, typename Real, class In>
__device__ __forceinline__ void stage_gather(const In& in, int t, cplx<Real> (&v)[N / T]) {
    static_assert(N % (R * T) == 0, "threads per line must divide N/R at every stage");
    constexpr int B = N / (R * T);   // butterflies per lane
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) v[b * R + r] = in(t + b * T + r * (N / R));
}

template <int N, int R, int NS, int DIR, int T, typename Real>
__device__ __forceinline__ void stage_compute(int t, const cplx<Real>* tw, cplx<Real> (&v)[N / T]) {
    constexpr int B = N / (R * T);
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int kk = (t + b * T) % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
            if (NS > 1) {
                cplx<Real> w = tw[(r * kk * (N / (NS * R))) % N];
                if (DIR > 0) w.y = -w.y;
                v[b * R + r] = cmul(v[b * R + r], w);
            }
        }
        Bfly<R, DIR, Real>::run(&v[b * R]);
    }
}

template <int N, int R, int NS, int DIR, int T, typename Real, class In>
__device__ __forceinline__ void stage_load(const In& in, int t, const cplx<Real>* tw, cplx<Real> (&v)[N / T]) {
    stage_gather<N, R, NS, DIR, T, Real>(in, t, v);
    stage_compute<N, R, NS, DIR, T, Real>(t, tw, v);
}

template <int N, int R, int NS, int T, typename Real, class Out>
__device__ __forceinline__ void stage_store(const Out& out, int t, const cplx<Real> (&v)[N / T]) {
    constexpr int B = N / (R * T);
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = t + b * T;
        const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
        for (int r = 0; r < R; ++r) out(base + r * NS, v[b * R + r]);
    }
}

template <int N, int S, int NS, int DIR, typename Real>
__device__ __forceinline__ void run_stages(cplx<Real>* line, int es, int t, bool active, const cplx<Real>* tw) {
    using P = FftPlan<N>;
    if constexpr (S < P::stages) {
        constexpr int R = P::radix[S];
        constexpr int T = P::threads;
        const LineRef<Real> ln{line, es};
        cplx<Real> v[N / T];
        if (active) stage_load<N, R, NS, DIR, T, Real>(ln, t, tw, v);
        __syncthreads();
        if (active) stage_store<N, R, NS, T, Real>(ln, t, v);
        __syncthreads();
        run_stages<N, S + 1, NS * R, DIR, Real>(line, es, t, active, tw);
    }
}

// Transform `nlines` lines in LDS: line l starts at base + l*ls, element j at
// + j*es (all in complex units).  Every thread of the workgroup must call this
// with the same arguments (it contains barriers).  tw: N-entry table in LDS.
// Lanes take the line index fastest: one access of a wave reads element
// j + r*N/R (or writes position k) of consecutive lines, which a line stride
// of N + 1 spreads over distinct banks.  With the lane index fastest instead,
// the T lanes of a line and the lines beside them meet on the same banks
// (element t + 8r of line l sits on bank l + t for N = 64, T = 8).
// lds_fft_lane: the same with the lane index `tid` of `nthr` lanes given by
// the caller.
template <int N, int DIR, typename Real>
__device__ __forceinline__ void lds_fft_lane(cplx<Real>* base, int nlines, int ls, int es, const cplx<Real>* tw, int tid,
                                             int nthr) {
    constexpr int T = FftPlan<N>::threads;
    const int lines_per_pass = nlines < nthr / T ? nlines : nthr / T;
    for (int l0 = 0; l0 < nlines; l0 += lines_per_pass) {
        const int l = l0 + tid % lines_per_pass;
        const int t = tid / lines_per_pass;
        const bool active = (t < T) && (l < nlines);
        cplx<Real>* line = base + (active ? l : 0) * ls;
        run_stages<N, 0, 1, DIR, Real>(line, es, t, active, tw);
    }
}

template <int N, int DIR, typename Real>
__device__ __forceinline__ void lds_fft(cplx<Real>* base, int nlines, int ls, int es, const cplx<Real>* tw) {
    lds_fft_lane<N, DIR, Real>(base, nlines, ls, es, tw, (int)threadIdx.x, (int)blockDim.x);
}

// ---------------------------------------------------------------------------
// Register-resident line transform with arbitrary endpoints.  Lane t (of T)
// of one line first gathers its stage-0 inputs (line_gather: plain loads,
// nothing waits on them), then line_finish runs the stages: stage 0 in
// registers, the middle stages through the line's LDS scratch `work`, and the
// last stage straight to `dst` (global or LDS).  With lanes of consecutive
// lines adjacent in a wave, the gather and the final store are coalesced
// across lines.  Barriers: every lane of the block must call line_finish.
//   SRC_LDS: the gathered source lives in LDS that `work` may overlap, so a
//            barrier separates all stage-0 reads from the first LDS write;
//   DST_LDS: dst is LDS (a barrier precedes the last-stage writes, which
//            then need one more barrier before anyone reads them).
// Stage 0 has no twiddles, so `tw` is first read after a barrier.
// ---------------------------------------------------------------------------
template <int N, typename Real, class Src>
__device__ __forceinline__ void line_gather(const Src& src, int t, bool active, cplx<Real> (&v)[N / FftPlan<N>::threads]) {
    using P = FftPlan<N>;
    if (active) stage_gather<N, P::radix[0], 1, -1, P::threads, Real>(src, t, v);
}

template <int N, int S, int NS, int DIR, typename Real, bool DST_LDS, class Dst>
__device__ __forceinline__ void line_rest(int t, bool active, const Dst& dst, const LineRef<Real>& work,
                                          const cplx<Real>* tw) {
    using P = FftPlan<N>;
    constexpr int R = P::radix[S];
    constexpr int T = P::threads;
    cplx<Real> v[N / T];
    if (active) stage_load<N, R, NS, DIR, T, Real>(work, t, tw, v);
    if constexpr (S + 1 == P::stages) {
        if constexpr (DST_LDS) __syncthreads();
        if (active) stage_store<N, R, NS, T, Real>(dst, t, v);
    } else {
        __syncthreads();
        if (active) stage_store<N, R, NS, T, Real>(work, t, v);
        __syncthreads();
        line_rest<N, S + 1, NS * R, DIR, Real, DST_LDS>(t, active, dst, work, tw);
    }
}

template <int N, int DIR, typename Real, bool SRC_LDS, bool DST_LDS, class Dst>
__device__ __forceinline__ void line_finish(cplx<Real> (&v)[N / FftPlan<N>::threads], int t, bool active,
                                            const Dst& dst, const LineRef<Real>& work, const cplx<Real>* tw) {
    using P = FftPlan<N>;
    constexpr int R = P::radix[0];
    constexpr int T = P::threads;
    if (active) stage_compute<N, R, 1, DIR, T, Real>(t, tw, v);
    if constexpr (P::stages == 1) {
        if constexpr (SRC_LDS && DST_LDS) __syncthreads();
        if (active) stage_store<N, R, 1, T, Real>(dst, t, v);
    } else {
        if constexpr (SRC_LDS) __syncthreads();
        if (active) stage_store<N, R, 1, T, Real>(work, t, v);
        __syncthreads();
        line_rest<N, 1, R, DIR, Real, DST_LDS>(t, active, dst, work, tw);
    }
}

}  // namespace mgk
