// gen_kernels.hip -- input generation on the device: counter-based random
// numbers (Philox4x32-10), the constructor's barrier and STZ-strain draws
// (kmc_solver.hpp:216-240, :1194-1215; BarrierGenerator :35-95; stz_mode_glass
// :1042-1074), and the correlated Gaussian field (material_generator.hpp:130-258).
//
// These draws follow the reference's distributions, not its std::mt19937
// stream: a trajectory started from them is a statistically equivalent one,
// not the reference's (which the host driver reproduces with its own
// generator).  The field filter is exact: given the same noise, the device
// field equals the reference's FFT-filtered one to rounding.
#include "common.hpp"
#include "gen_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace mgk {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011): ten rounds of two 32x32 -> 64-bit multiplies on a
// 128-bit counter under a 64-bit key bumped by the Weyl constants.
// ---------------------------------------------------------------------------
struct Words4 {
    unsigned w[4];
};

__host__ __device__ inline Words4 philox4x32_10(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0,
                                                unsigned k1) {
    const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        const unsigned long long p0 = (unsigned long long)M0 * c0, p1 = (unsigned long long)M1 * c2;
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
        const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return Words4{{c0, c1, c2, c3}};
}

// A uniform in (0, 1) from one word (never 0: logarithms stay finite).
__host__ __device__ inline double unit_open(unsigned w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

// Stream of draws for item i: counter (i lo, i hi, round, purpose), key = seed.
struct DrawStream {
    unsigned i0, i1, purpose, k0, k1;
    unsigned round = 0;
    Words4 buf{};
    int used = 4;
    __device__ DrawStream(unsigned long long item, unsigned purpose_, unsigned long long seed)
        : i0((unsigned)item), i1((unsigned)(item >> 32)), purpose(purpose_), k0((unsigned)seed),
          k1((unsigned)(seed >> 32)) {}
    __device__ unsigned word() {
        if (used == 4) {
            buf = philox4x32_10(i0, i1, round++, purpose, k0, k1);
            used = 0;
        }
        return buf.w[used++];
    }
    __device__ double uniform() { return unit_open(word()); }
    // box_muller_normal (material_generator.hpp:121-128) on open uniforms
    __device__ double normal() {
        const double u1 = uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
};

// ---------------------------------------------------------------------------
// BarrierGenerator::generate (kmc_solver.hpp:35-95) on a device stream.
// ---------------------------------------------------------------------------
__device__ double modified_rayleigh(const DrawParams& p, DrawStream& s) {
    const double a = fmax(0.0, p.b_mean - 6.0 * p.b_std);
    const double b = p.b_mean + 6.0 * p.b_std;
    const double x_max = (p.b_mean + sqrt(p.b_mean * p.b_mean + 4.0 * p.b_std * p.b_std)) / 2.0;
    const double c_max = x_max * exp(-(x_max - p.b_mean) * (x_max - p.b_mean) / (2.0 * p.b_std * p.b_std));
    for (int k = 0; k < kMaxRejections; ++k) {
        const double x = a + (b - a) * s.uniform();
        const double u = c_max * s.uniform();
        if (u < x * exp(-(x - p.b_mean) * (x - p.b_mean) / (2.0 * p.b_std * p.b_std))) return x;
    }
    return x_max;   // kMaxRejections rounds rejected (probability < 1e-90)
}

__device__ double barrier(const DrawParams& p, DrawStream& s) {
    double v = 0.0;
    switch (p.barrier_type) {
        case 0: v = p.b_mean + p.b_std * s.normal(); break;                                   // gaussian
        case 1: v = p.b_loc + p.b_scale * sqrt(-2.0 * log(1.0 - s.uniform())); break;        // rayleigh
        case 2: v = modified_rayleigh(p, s); break;
        case 3:
            v = (s.uniform() < p.b_ratio) ? -p.b_epsilon * log(1.0 - s.uniform()) : modified_rayleigh(p, s);
            break;
        default: v = 0.0;
    }
    if (p.b_min_cutoff >= 0.0 && v < p.b_min_cutoff) v = p.b_min_cutoff;
    if (p.b_max_cutoff >= 0.0 && v > p.b_max_cutoff) v = p.b_max_cutoff;
    return v;
}

// One STZ strain (symmetric components): the planar mode gxx, gxy ~
// 0.5 gamma0 N(0, 1) (:232-237, :1042-1052) or, in 3D with use_3d_barriers, a
// random symmetric deviatoric of Frobenius norm 1/sqrt(2) scaled by
// gamma0 sqrt(-2 ln u) (:1054-1074).
template <int C>
__device__ void stz_mode(const DrawParams& p, DrawStream& s, double* c) {
    for (int k = 0; k < C; ++k) c[k] = 0.0;
    if (C == 6 && p.glass_3d) {
        double m[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) m[i][j] = s.normal();
        double g[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) g[i][j] = 0.5 * (m[i][j] + m[j][i]);
        const double t = (g[0][0] + g[1][1] + g[2][2]) / 3.0;
        for (int i = 0; i < 3; ++i) g[i][i] -= t;
        double nsq = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) nsq += g[i][j] * g[i][j];
        const double nrm = sqrt(nsq);
        if (nrm < 1e-12) return;
        const double u = fmax(s.uniform(), 1e-15);
        const double f = p.gamma0 * sqrt(-2.0 * log(u)) / (sqrt(2.0) * nrm);
        const int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
        for (int k = 0; k < 6; ++k) c[k] = g[ii[k]][jj[k]] * f;
        return;
    }
    const double gxx = 0.5 * p.gamma0 * s.normal();
    const double gxy = 0.5 * p.gamma0 * s.normal();
    c[0] = gxx;
    c[1] = -gxx;
    c[C == 3 ? 2 : 3] = gxy;
}

// Q0 and the per-(voxel, mode) catalog, one lane per mode: [C][N*M] catalog.
template <int C>
__global__ __launch_bounds__(256) void k_draw_modes(double* __restrict__ Q0, double* __restrict__ cat, long nmodes,
                                                    const DrawParams p, unsigned long long seed) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nmodes) return;
    DrawStream sb((unsigned long long)i, 1u, seed);
    Q0[i] = barrier(p, sb);
    if (cat) {
        DrawStream sc((unsigned long long)i, 2u, seed);
        double c[C];
        stz_mode<C>(p, sc, c);
        for (int k = 0; k < C; ++k) cat[k * nmodes + i] = c[k];
    }
}

// ---------------------------------------------------------------------------
// Correlated field.  The reference filters white noise with the Fourier
// multiplier exp(-2 pi^2 sigma^2 |f|^2), f = fftfreq(n) per axis
// (material_generator.hpp:160-176, :196-219).  The multiplier is separable, so
// the filter is three circular convolutions along the axes with the periodic
// kernels h_a = IDFT_n(exp(-2 pi^2 sigma^2 f^2)), computed on the host.
// ---------------------------------------------------------------------------
__global__ void k_normals(double* __restrict__ out, long n, unsigned long long seed) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DrawStream s((unsigned long long)i, 3u, seed);
    out[i] = s.normal();
}

// out = h (*) in along the axis of length n and stride st; lines are the other
// axes.  One block per line, the line staged in LDS.
__global__ __launch_bounds__(256) void k_conv_axis(const double* __restrict__ in, double* __restrict__ out,
                                                   const double* __restrict__ h, int n, long st, long inner,
                                                   long nlines) {
    extern __shared__ double line[];
    double* hk = line + n;
    const long L = blockIdx.x;
    if (L >= nlines) return;
    const long base = (L / inner) * inner * n + (L % inner);   // line start: outer * (n * inner) + inner index
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        line[j] = in[base + (long)j * st];
        hk[j] = h[j];
    }
    __syncthreads();
    for (int x = threadIdx.x; x < n; x += blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) {
            int src = x - j;
            if (src < 0) src += n;
            acc += hk[j] * line[src];
        }
        out[base + (long)x * st] = acc;
    }
}

// Block partials of sum(v) (pass 0) or sum((v - mean)^2) (pass 1).
__global__ __launch_bounds__(256) void k_moments(const double* __restrict__ v, long n, const double* __restrict__ mean,
                                                 double* __restrict__ part, int pass) {
    __shared__ double red[256];
    double acc = 0.0;
    const double m = pass ? mean[0] : 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double d = v[i] - m;
        acc += pass ? d * d : v[i];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// stat[0] = mean or stat[1] = population std (1 if < 1e-15), from the partials
__global__ void k_moments_final(const double* __restrict__ part, int nb, long n, double* __restrict__ stat, int pass) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    if (pass == 0) stat[0] = s / (double)n;
    else {
        double sd = sqrt(s / (double)n);
        if (sd < 1e-15) sd = 1.0;
        stat[1] = sd;
    }
}

// field = mean_t + std_t * (v - mean) / sd, then the clips (:222-255)
__global__ __launch_bounds__(256) void k_standardize(double* __restrict__ v, long n, const double* __restrict__ stat,
                                                     FieldParams f) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = f.mean + f.std * ((v[i] - stat[0]) / stat[1]);
    if (f.has_min && x < f.clip_min) x = f.clip_min;
    if (f.has_max && x > f.clip_max) x = f.clip_max;
    v[i] = x;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void gen_enqueue_modes(double* Q0, double* cat, long nmodes, int C, const DrawParams& p, unsigned long long seed,
                       hipStream_t s) {
    const int nb = (int)((nmodes + 255) / 256);
    if (C == 3) hipLaunchKernelGGL((k_draw_modes<3>), dim3(nb), dim3(256), 0, s, Q0, cat, nmodes, p, seed);
    else hipLaunchKernelGGL((k_draw_modes<6>), dim3(nb), dim3(256), 0, s, Q0, cat, nmodes, p, seed);
    MGK_HIP_CHECK(hipGetLastError());
}

// Periodic kernel of the multiplier exp(c f^2), f = fftfreq(n): h[x] =
// (1/n) sum_k g_k cos(2 pi k x / n) (g is even in f, so h is real).
static std::vector<double> filter_kernel(int n, double sigma) {
    std::vector<double> h(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    const long double c = -2.0L * pi * pi * (long double)sigma * (long double)sigma;
    std::vector<long double> g(n);
    const int lim = (n + 1) / 2;
    for (int k = 0; k < n; ++k) {
        const long double f = (long double)(k < lim ? k : k - n) / (long double)n;
        g[k] = expl(c * f * f);
    }
    for (int x = 0; x < n; ++x) {
        long double acc = 0.0L;
        for (int k = 0; k < n; ++k) acc += g[k] * cosl(2.0L * pi * (long double)((long)k * x % n) / (long double)n);
        h[x] = (double)(acc / (long double)n);
    }
    return h;
}

void gen_correlated_field(const int* n, int dims, const FieldParams& f, unsigned long long seed, const double* noise,
                          double* out) {
    long N = 1;
    for (int a = 0; a < dims; ++a) N *= n[a];
    double *a = nullptr, *b = nullptr, *h = nullptr, *part = nullptr, *stat = nullptr;
    const int nb = 256;
    auto cleanup = [&]() {
        for (double* p : {a, b, h, part, stat})
            if (p) (void)hipFree(p);
    };
    try {
        MGK_HIP_CHECK(hipMalloc(&a, N * sizeof(double)));
        MGK_HIP_CHECK(hipMalloc(&b, N * sizeof(double)));
        int nmax = 1;
        for (int k = 0; k < dims; ++k) nmax = std::max(nmax, n[k]);
        MGK_HIP_CHECK(hipMalloc(&h, nmax * sizeof(double)));
        MGK_HIP_CHECK(hipMalloc(&part, nb * sizeof(double)));
        MGK_HIP_CHECK(hipMalloc(&stat, 2 * sizeof(double)));
        if (noise) MGK_HIP_CHECK(hipMemcpy(a, noise, N * sizeof(double), hipMemcpyHostToDevice));
        else hipLaunchKernelGGL(k_normals, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, 0, a, N, seed);
        // axes from the last (contiguous) to the first; each pass ping-pongs a <-> b
        long inner = 1;
        for (int ax = dims - 1; ax >= 0; --ax) {
            const std::vector<double> hk = filter_kernel(n[ax], f.corr);
            MGK_HIP_CHECK(hipMemcpy(h, hk.data(), hk.size() * sizeof(double), hipMemcpyHostToDevice));
            const long nlines = N / n[ax];
            hipLaunchKernelGGL(k_conv_axis, dim3((unsigned)nlines), dim3(std::min(256, ((n[ax] + 63) / 64) * 64)),
                               2 * n[ax] * sizeof(double), 0, a, b, h, n[ax], inner, inner, nlines);
            MGK_HIP_CHECK(hipGetLastError());
            std::swap(a, b);
            inner *= n[ax];
        }
        hipLaunchKernelGGL(k_moments, dim3(nb), dim3(256), 0, 0, a, N, stat, part, 0);
        hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(64), 0, 0, part, nb, N, stat, 0);
        hipLaunchKernelGGL(k_moments, dim3(nb), dim3(256), 0, 0, a, N, stat, part, 1);
        hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(64), 0, 0, part, nb, N, stat, 1);
        hipLaunchKernelGGL(k_standardize, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, 0, a, N, stat, f);
        MGK_HIP_CHECK(hipGetLastError());
        MGK_HIP_CHECK(hipMemcpy(out, a, N * sizeof(double), hipMemcpyDeviceToHost));
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

// ---------------------------------------------------------------------------
// The correlated field without allocation: the same kernels, with every
// buffer beyond the caller's noise and output in `scratch`, a file-scope
// device array (one copy per GPU, held until the process ends):
//   [0, kNoAllocMaxVoxels)            the ping-pong field buffer
//   then 3 x kNoAllocMaxAxis          the filter taps of each axis
//   then kNoAllocParts + 2            the moment partials, mean and deviation
// The taps are computed on the device (k_filter_taps) instead of on the host.
// Calls on one GPU share the array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int kNoAllocParts = 256;
constexpr size_t kScratchDoubles = (size_t)kNoAllocMaxVoxels + 3 * (size_t)kNoAllocMaxAxis + kNoAllocParts + 2;
__device__ __align__(256) double scratch[kScratchDoubles];

// h[x] = (1/n) sum_k exp(c f_k^2) cos(2 pi ((k x) mod n) / n), f = fftfreq(n):
// filter_kernel's taps, summed with compensation in FP64
__global__ __launch_bounds__(256) void k_filter_taps(double* __restrict__ h, int n, double corr) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const double pi = 3.141592653589793;
    const double c = -2.0 * pi * pi * corr * corr;
    const int lim = (n + 1) / 2;
    double acc = 0.0, comp = 0.0;
    for (int k = 0; k < n; ++k) {
        const double f = (double)(k < lim ? k : k - n) / (double)n;
        const double term = exp(c * f * f) * cospi(2.0 * (double)(((long)k * x) % n) / (double)n) - comp;
        const double t = acc + term;
        comp = (t - acc) - term;
        acc = t;
    }
    h[x] = acc / (double)n;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

// The device address of `scratch` on the current GPU, looked up once per GPU
// (the array's host-side name is not a device address).
double* scratch_address() {
    int dev = 0;
    MGK_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw ArgError("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        MGK_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw HipError("hipGetSymbolAddress(scratch) returned no address");
        g_scratch[dev] = static_cast<double*>(ptr);
    }
    return g_scratch[dev];
}

void gen_correlated_field_noalloc(const int* n, int dims, const FieldParams& f, unsigned long long seed,
                                  const double* noise, double* out, hipStream_t s) {
    long N = 1;
    for (int a = 0; a < dims; ++a) {
        if (n[a] > kNoAllocMaxAxis)
            throw ArgError("axis of length " + std::to_string(n[a]) + " over the no-allocation limit of " +
                           std::to_string(kNoAllocMaxAxis));
        N *= n[a];
    }
    if (N > kNoAllocMaxVoxels)
        throw ArgError("grid of " + std::to_string(N) + " voxels over the no-allocation limit of " +
                       std::to_string(kNoAllocMaxVoxels));
    double* sc = scratch_address();
    double* y = sc;                                   // ping-pong partner of `out`
    double* taps = sc + kNoAllocMaxVoxels;            // 3 x kNoAllocMaxAxis
    double* part = taps + 3 * (size_t)kNoAllocMaxAxis;
    double* stat = part + kNoAllocParts;
    for (int a = 0; a < dims; ++a)
        hipLaunchKernelGGL(k_filter_taps, dim3((unsigned)((n[a] + 255) / 256)), dim3(256), 0, s,
                           taps + (size_t)a * kNoAllocMaxAxis, n[a], f.corr);
    // dims passes alternate between `out` and y and must end in `out`
    const double* src;
    double* dst;
    if (noise) {
        src = noise;
        dst = (dims % 2) ? out : y;
    } else {
        double* start = (dims % 2) ? y : out;
        hipLaunchKernelGGL(k_normals, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, start, N, seed);
        src = start;
        dst = (start == out) ? y : out;
    }
    long inner = 1;
    for (int ax = dims - 1; ax >= 0; --ax) {
        const long nlines = N / n[ax];
        hipLaunchKernelGGL(k_conv_axis, dim3((unsigned)nlines), dim3(std::min(256, ((n[ax] + 63) / 64) * 64)),
                           2 * n[ax] * sizeof(double), s, src, dst, taps + (size_t)ax * kNoAllocMaxAxis, n[ax], inner,
                           inner, nlines);
        src = dst;
        dst = (dst == out) ? y : out;
        inner *= n[ax];
    }
    hipLaunchKernelGGL(k_moments, dim3(kNoAllocParts), dim3(256), 0, s, out, N, stat, part, 0);
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(64), 0, s, part, kNoAllocParts, N, stat, 0);
    hipLaunchKernelGGL(k_moments, dim3(kNoAllocParts), dim3(256), 0, s, out, N, stat, part, 1);
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(64), 0, s, part, kNoAllocParts, N, stat, 1);
    hipLaunchKernelGGL(k_standardize, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, out, N, stat, f);
    MGK_HIP_CHECK(hipGetLastError());
}

}  // namespace mgk
