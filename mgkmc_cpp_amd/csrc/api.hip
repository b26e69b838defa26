// api.hip -- the C ABI of include/mgkmc_gpu.h: context lifecycle, the
// Lippmann-Schwinger solve driver (batched enqueue, device-side convergence
// flag, one host sync per batch), device-resident KMC state and the per-event
// primitives.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/mgkmc_gpu.h"
#include "common.hpp"
#include "gen_kernels.hpp"
#include "kmc_kernels.hpp"
#include "ls_kernels.hpp"

using namespace mgk;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

// Calls that allocate, free or synchronise the whole device (hipMalloc,
// hipFree, hipDeviceSynchronize, synchronous copies) must not overlap a
// stream capture in another thread: they would invalidate it.  Entry points
// that make them, and every graph capture, hold this lock.
std::recursive_mutex g_device_mu;

// A persistent CG cycle's blocks wait for each other at grid barriers, so all
// of them must be resident at once.  It is an ordinary launch (a cooperative
// launch cost about 22 us per event at 64^3 in stream synchronisation around
// it); the device holds its blocks (checked once with the occupancy API), and
// this lock keeps a second cycle from another context from competing for the
// same slots: a thread holds it from a cycle's launch until its stream has
// finished the round.  Kernels of other streams do not wait on a cycle, so
// blocks they hold free up; a barrier that still waits past its poll limit
// fails the solve with an error instead of hanging.
std::mutex g_cycle_mu;

// Flips of a cascade round uploaded and applied per launch (mgk_apply_flips).
constexpr int kFlipChunk = 64;

template <typename F> int guarded(F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return fail(MGK_E_HIP, e.what);
    } catch (const ArgError& e) {
        return fail(MGK_E_ARG, e.what);
    } catch (const UnsupportedError& e) {
        return fail(MGK_E_UNSUPPORTED, e.what);
    } catch (const std::exception& e) {
        return fail(MGK_E_HIP, e.what());
    }
}

template <typename T> T* dalloc(size_t n) {
    void* p = nullptr;
    MGK_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T*>(p);
}
template <typename T> T* halloc(size_t n) {
    void* p = nullptr;
    MGK_HIP_CHECK(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
    return static_cast<T*>(p);
}

// exp(-2 pi i k / n), k < n, rounded from long double
template <typename R> std::vector<R> twiddle_table(int n) {
    std::vector<R> t(2 * (size_t)n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n; ++k) {
        long double a = 2.0L * pi * (long double)k / (long double)n;
        t[2 * k] = (R)cosl(a);
        t[2 * k + 1] = (R)(-sinl(a));
    }
    return t;
}

// 64-bit FNV-1a over a byte range, continuing from h.
unsigned long long fnv1a(const void* data, size_t n, unsigned long long h = 1469598103934665603ull) {
    const unsigned char* b = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

void lame(int dims, int plane_mode, double E, double nu, double& lam, double& mu) {
    // compute_lame_2d / compute_lame_3d (linear_elastic_solver.hpp:36-45, :315-319)
    mu = E / (2.0 * (1.0 + nu));
    if (dims == 2 && plane_mode == MGK_PLANE_STRESS) lam = E * nu / (1.0 - nu * nu);
    else lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
}

}  // namespace

struct mgk_ctx {
    Geom g;
    double pixel = 1.0;
    int plane_mode = MGK_PLANE_STRAIN;
    int precision = MGK_FP64;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // block-shape variants and CG pipeline chosen by the autotuner
    LsTune tune{};
    int use_cycle = 1;       // the persistent CG cycle where the grid supports it (else the plane passes)
    // in-situ timing of the persistent CG cycle launches (mgk_cycle_timer)
    hipEvent_t cyc_ev0 = nullptr, cyc_ev1 = nullptr;
    bool cyc_timing = false;
    double cyc_ms = 0.0;
    long long cyc_launches = 0, cyc_steps = 0;
    int device = 0;

    // material
    double *E = nullptr, *nu = nullptr, *lam = nullptr, *mu = nullptr;
    float *dlam32 = nullptr, *dmu32 = nullptr;
    double lam0 = 0, mu0 = 0, mean_mu = 0;
    std::vector<double> h_mu;
    unsigned long long material_hash = 0;   // FNV-1a of E and nu, checked when a state is loaded

    // strain state (symmetric SoA) + uniform antisymmetric part on the host
    double *eps = nullptr, *sig = nullptr, *epsp = nullptr;
    bool epsp_set = false;
    double W[3] = {0, 0, 0};   // 2D: w_xy; 3D: w_xy, w_xz, w_yz

    // spectra and FP32 refinement buffers
    void* spec = nullptr;
    float *cg = nullptr, *corr = nullptr;   // CG vectors r, w, p, s (4 x C x N) and iterate x
    double* partials = nullptr;
    Ctl* ctl = nullptr;
    Ctl* hctl = nullptr;
    void *tw64[3] = {nullptr, nullptr, nullptr}, *tw32[3] = {nullptr, nullptr, nullptr};

    // staging for AoS <-> SoA transfers
    double* dstage = nullptr;
    size_t dstage_n = 0;
    double* hstage = nullptr;
    size_t hstage_n = 0;

    // statistics
    int last_iters = 0, last_inner = 0, predicted = 4;
    int inner_per_outer = 12;          // CG steps enqueued for a solve's first refinement cycle
    int inner_hist[4] = {0, 0, 0, 0};
    unsigned inner_pos = 0;
    int inner_per_outer2 = 4;          // ... and for each later cycle
    int inner_hist2[4] = {0, 0, 0, 0};
    unsigned inner_pos2 = 0;
    bool mean_valid = false;
    double mean_eps[6] = {0, 0, 0, 0, 0, 0};

    // KMC
    bool kmc = false;
    mgk_kmc_params kp{};
    KmcParams K{};
    KmcCtl* kctl = nullptr;
    KmcCtl* hkctl = nullptr;
    double* kstage = nullptr;
    double* hkstage = nullptr;
    // a cascade round's flips, uploaded kFlipChunk at a time (mgk_apply_flips)
    double* fstage = nullptr;
    double* hfstage = nullptr;
    int* fidx = nullptr;
    int* hfidx = nullptr;
    int* unst_idx = nullptr;
    double* unst_q = nullptr;
    std::vector<double> shared_cat;   // [C][M] host copy of a shared catalog
    bool kctl_pending = false;        // a device-side selection not yet copied back

    // parameter blocks of this context's launches in device memory (LS slot
    // 0: solves, slot 1: pass timing), uploaded from pinned copies when they
    // change; one slot of event arguments
    LsParams* d_lsp = nullptr;
    LsParams* h_lsp = nullptr;
    bool lsp_up[2] = {false, false};
    KmcParams* d_kp = nullptr;
    KmcParams* h_kp = nullptr;
    bool kp_up = false;
    EventArgs* d_ea = nullptr;
    EventArgs* h_ea = nullptr;

    // fast patching: kernels [nb][C][W], missing means [nb][C], sigma_macro_unit [C]
    double *pk = nullptr, *pkm = nullptr, *psu = nullptr;
    PatchParams pp{};
    bool patch_ready = false, unit_ready = false;

    // device bytes this context holds (every allocation goes through dal)
    size_t dev_bytes = 0;
    size_t kmc_bytes = 0;   // the part held by the KMC state
    size_t cat_cap = 0;     // doubles of catalog storage
    template <typename T> T* dal(size_t n, bool kmc_state = false) {
        T* p = dalloc<T>(n);
        const size_t b = std::max<size_t>(n, 1) * sizeof(T);
        dev_bytes += b;
        if (kmc_state) kmc_bytes += b;
        return p;
    }
    // catalog storage of at least n doubles: M*C for a catalog every voxel
    // shares, C*N*M for a per-voxel one (allocated when first needed)
    void ensure_cat(size_t n) {
        if (n <= cat_cap) return;
        double* p = dal<double>(n, true);
        if (K.cat) {
            MGK_HIP_CHECK(hipFree(K.cat));
            dev_bytes -= std::max<size_t>(cat_cap, 1) * sizeof(double);
            kmc_bytes -= std::max<size_t>(cat_cap, 1) * sizeof(double);
        }
        K.cat = p;
        cat_cap = n;
    }

    ~mgk_ctx() { release(); }

    void release() {
        auto f = [](void* p) { if (p) (void)hipFree(p); };
        auto fh = [](void* p) { if (p) (void)hipHostFree(p); };
        f(E); f(nu); f(lam); f(mu); f(dlam32); f(dmu32); f(eps); f(sig); f(epsp); f(spec); f(cg); f(corr);
        f(partials); f(ctl); f(dstage);
        for (int a = 0; a < 3; ++a) { f(tw64[a]); f(tw32[a]); }
        fh(hctl); fh(hstage);
        f(K.Q); f(K.Q0); f(K.cat); f(K.soft); f(K.last_t); f(K.prev_dir); f(K.bsum); f(K.bincl);
        f(K.bunst); f(K.bunst_off); f(K.binc); f(kstage); f(unst_idx); f(unst_q);   // kctl: inside ctl's block
        f(fstage); f(fidx);
        fh(hkstage); fh(hfstage); fh(hfidx);
        f(d_lsp); f(d_kp); f(d_ea); f(pk); f(pkm); f(psu);
        fh(h_lsp); fh(h_kp); fh(h_ea);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (cyc_ev0) (void)hipEventDestroy(cyc_ev0);
        if (cyc_ev1) (void)hipEventDestroy(cyc_ev1);
        if (stream) (void)hipStreamDestroy(stream);
        E = nullptr;
        stream = nullptr;
        ev0 = ev1 = nullptr;
        cyc_ev0 = cyc_ev1 = nullptr;
    }

    int D() const { return g.dims; }
    int full() const { return g.dims * g.dims; }

    double* stage_dev(size_t n) {
        if (n > dstage_n) {
            if (dstage) {
                MGK_HIP_CHECK(hipFree(dstage));
                dev_bytes -= std::max<size_t>(dstage_n, 1) * sizeof(double);
            }
            dstage = dal<double>(n);
            dstage_n = n;
        }
        return dstage;
    }
    double* stage_host(size_t n) {
        if (n > hstage_n) {
            if (hstage) MGK_HIP_CHECK(hipHostFree(hstage));
            hstage = halloc<double>(n);
            hstage_n = n;
        }
        return hstage;
    }

    LsParams ls_params() const {
        LsParams p;
        std::memset(&p, 0, sizeof(p));   // compared byte-wise against the uploaded copy
        p.g = g;
        p.eps = eps;
        p.epsp = epsp_set ? epsp : nullptr;
        p.lam = lam;
        p.mu = mu;
        p.sig = sig;
        p.lam0 = lam0;
        p.mu0 = mu0;
        p.spec64 = spec;
        p.spec32 = spec;
        const long CN = (long)g.C * g.N;
        p.r32 = (precision == MGK_FP32_REFINE) ? cg : nullptr;
        p.w32 = cg + CN;
        p.p32 = cg + 2 * CN;
        p.s32 = cg + 3 * CN;
        p.mandel = ls_plane_path(g) ? 1 : 0;
        p.corr = (precision == MGK_FP32_REFINE) ? corr : nullptr;
        p.dlam32 = dlam32;
        p.dmu32 = dmu32;
        for (int a = 0; a < 3; ++a) {
            (&p.tw_f64)[a] = tw64[a];
            (&p.tw_f32)[a] = tw32[a];
        }
        p.ctl = ctl;
        p.partials = partials;
        p.gbar = (precision == MGK_FP32_REFINE && use_cycle && ls_cycle_supported(g))
                     ? reinterpret_cast<unsigned*>(partials + ls_gbar_offset(g)) : nullptr;
        p.r0_cart = (precision == MGK_FP32_REFINE && ls_plane64_path(g)) ? 1 : 0;
        return p;
    }

    // This context's LS parameter block in device memory (slot 0), or a
    // variant of it (slot 1), re-uploaded only when it changed.
    LsLaunch lsl(int slot = 0, const LsParams* custom = nullptr) {
        const LsParams p = custom ? *custom : ls_params();
        if (!lsp_up[slot] || std::memcmp(&h_lsp[slot], &p, sizeof(LsParams)) != 0) {
            MGK_HIP_CHECK(hipStreamSynchronize(stream));   // the pinned copy may still feed an earlier upload
            h_lsp[slot] = p;
            MGK_HIP_CHECK(hipMemcpyAsync(d_lsp + slot, h_lsp + slot, sizeof(LsParams), hipMemcpyHostToDevice, stream));
            lsp_up[slot] = true;
        }
        LsLaunch L{g, d_lsp + slot, 1, h_lsp[slot], false, h_lsp[slot].mandel != 0};
        L.plane64 = ls_plane64_path(g);
        L.tune = tune;
        return L;
    }
    KmcLaunch kml() {
        if (!kp_up || std::memcmp(h_kp, &K, sizeof(KmcParams)) != 0) {
            MGK_HIP_CHECK(hipStreamSynchronize(stream));
            *h_kp = K;
            MGK_HIP_CHECK(hipMemcpyAsync(d_kp, h_kp, sizeof(KmcParams), hipMemcpyHostToDevice, stream));
            kp_up = true;
        }
        return KmcLaunch{g, kp.M, d_kp, 1, K, false};
    }

    // Symmetric / antisymmetric split of a column-major D x D matrix.
    void split(const double* a, double* sym, double* skew) const {
        const int Dd = D();
        for (int i = 0; i < Dd; ++i)
            for (int j = i; j < Dd; ++j) {
                const int c = comp_of(Dd, i, j);
                if (i == j) sym[c] = a[i + Dd * i];
                else {
                    sym[c] = 0.5 * (a[i + Dd * j] + a[j + Dd * i]);
                    const int w = (Dd == 2) ? 0 : (c - 3);
                    skew[w] = 0.5 * (a[i + Dd * j] - a[j + Dd * i]);
                }
            }
    }
    void join(const double* sym, const double* skew, double* a) const {
        const int Dd = D();
        for (int i = 0; i < Dd; ++i)
            for (int j = 0; j < Dd; ++j) {
                const int c = comp_of(Dd, i, j);
                if (i == j) a[i + Dd * j] = sym[c];
                else {
                    const int w = (Dd == 2) ? 0 : (comp_of(Dd, i, j) - 3);
                    a[i + Dd * j] = sym[c] + (i < j ? skew[w] : -skew[w]);
                }
            }
    }
    int nskew() const { return D() == 2 ? 1 : 3; }

    // AoS (N x D x D, column-major) host array -> symmetric SoA device field.
    // Returns the per-voxel antisymmetric parts through skew_out (may be null).
    void upload_tensor(const double* host, double* dev, std::vector<double>* skew_out) {
        const long N = g.N;
        const int C = g.C, F = full(), ns = nskew();
        double* h = stage_host((size_t)N * C);
        if (skew_out) skew_out->assign((size_t)N * ns, 0.0);
        double sym[6], sk[3] = {0, 0, 0};
        for (long v = 0; v < N; ++v) {
            split(host + v * F, sym, sk);
            for (int c = 0; c < C; ++c) h[(size_t)c * N + v] = sym[c];
            if (skew_out)
                for (int w = 0; w < ns; ++w) (*skew_out)[(size_t)v * ns + w] = sk[w];
        }
        MGK_HIP_CHECK(hipMemcpyAsync(dev, h, sizeof(double) * N * C, hipMemcpyHostToDevice, stream));
        MGK_HIP_CHECK(hipStreamSynchronize(stream));
    }
    // symmetric SoA device field -> AoS host array with per-voxel skew(v) added.
    template <typename SkewFn> void download_tensor(const double* dev, double* host, SkewFn skew_of) {
        const long N = g.N;
        const int C = g.C, F = full(), ns = nskew();
        double* h = stage_host((size_t)N * C);
        MGK_HIP_CHECK(hipMemcpyAsync(h, dev, sizeof(double) * N * C, hipMemcpyDeviceToHost, stream));
        MGK_HIP_CHECK(hipStreamSynchronize(stream));
        double sym[6], sk[3];
        for (long v = 0; v < N; ++v) {
            for (int c = 0; c < C; ++c) sym[c] = h[(size_t)c * N + v];
            skew_of(v, sk);
            (void)ns;
            join(sym, sk, host + v * F);
        }
    }

    // The LS solve on the device-resident fields (spectral_solver_*, warm
    // start when `warm`).  eps_bar_sym is the imposed strain's symmetric part,
    // norm_extra the ||eps||^2 contribution of its uniform antisymmetric part.
    // Returns the FP64 iteration count.
    int solve(const double* eps_bar_sym, double norm_extra, bool warm, int max_iter, double tol, double* diff_out) {
        Ctl& c = *hctl;
        init_ctl(max_iter, tol);
        c.norm_extra = norm_extra;
        for (int k = 0; k < g.C; ++k) c.eps_bar[k] = eps_bar_sym[k];
        const int start = start_mode(warm);
        if (start == LS_START_KEEP) {   // the field is unchanged since the last solve: its mean is known
            for (int k = 0; k < g.C; ++k) c.shift[k] = eps_bar_sym[k] - mean_eps[k];
            c.keep_shift = 1;
        }
        const LsLaunch L = lsl();
        MGK_HIP_CHECK(hipMemcpyAsync(ctl, hctl, sizeof(Ctl), hipMemcpyHostToDevice, stream));
        ls_enqueue_start(L, start, stream);
        return run_batches(warm, max_iter, tol, diff_out, false);
    }

    void init_ctl(int max_iter, double tol) {
        std::memset(hctl, 0, sizeof(Ctl));
        hctl->max_iter = max_iter;
        hctl->tol = tol;
        for (int k = 0; k < g.C; ++k) hctl->mean_eps[k] = mean_eps[k];
    }
    int start_mode(bool warm) const {
        return (warm && mean_valid) ? LS_START_KEEP : (warm ? LS_START_MEAN : LS_START_FILL);
    }

    // Iteration batches until the control block says done, the final pass
    // and the read-back: one host sync per solve when the CG-step prediction
    // holds.  `event`: the solve sits behind a device-decided KMC event
    // attempt whose control block comes back with it; an attempt without an
    // event left the solve marked done before its first iteration, and the
    // solver statistics are then left alone.
    int run_batches(bool warm, int max_iter, double tol, double* diff_out, bool event) {
        const LsLaunch p = lsl();
        const bool refine = (precision == MGK_FP32_REFINE);
        int enqueued = 0;   // FP64 iterations (refinement cycles) enqueued
        bool final_done = false;
        int inner_before = 0;   // CG steps counted before this batch
        std::unique_lock<std::mutex> cycle_lock(g_cycle_mu, std::defer_lock);   // held while a cycle is in flight
        if (max_iter > 0) {
            int batch = std::min(max_iter, std::max(refine ? 1 : 2, predicted));
            while (true) {
                // a batch with a single cycle launch is timed when asked
                const bool timed = cyc_timing && refine && p.h.gbar && batch == 1;
                for (int b = 0; b < batch; ++b) {
                    const int k = enqueued + b;
                    const int flags = ((k == 0 && warm) ? LS_FLAG_SHIFT : 0) | ((refine && k > 0) ? LS_FLAG_CORR : 0);
                    ls_enqueue_iteration(p, LS_FULL, flags, stream);
                    if (refine) {
                        const int J = (k == 0) ? inner_per_outer : inner_per_outer2;
                        if (p.h.gbar) {
                            if (!cycle_lock.owns_lock()) cycle_lock.lock();
                            if (timed) MGK_HIP_CHECK(hipEventRecord(cyc_ev0, stream));
                            ls_enqueue_cycle(p, J, stream);   // the whole cycle in one launch
                            if (timed) MGK_HIP_CHECK(hipEventRecord(cyc_ev1, stream));
                        } else {
                            const int sp = (k == 0) ? spec_from() : spec_from2();
                            for (int j = 0; j < J; ++j)
                                ls_enqueue_iteration(p, LS_CG, cg_flags(j, sp, J), stream);
                        }
                    }
                }
                enqueued += batch;
                // the final pass runs with the batch: one host sync per solve
                // when the prediction holds; it is repeated if more iterations follow
                ls_enqueue_final(p, stream);
                // the KMC control block sits right after the LS one (CtlPair)
                MGK_HIP_CHECK(hipMemcpyAsync(hctl, ctl, kctl_pending ? sizeof(CtlPair) : sizeof(Ctl),
                                             hipMemcpyDeviceToHost, stream));
                kctl_pending = false;
                MGK_HIP_CHECK(hipStreamSynchronize(stream));
                if (cycle_lock.owns_lock()) cycle_lock.unlock();
                if (hctl->cg_err) throw HipError("persistent CG cycle: grid barrier timed out");
                if (timed && hctl->inner > inner_before) {   // the cycle ran (not converged before it)
                    float ms = 0.f;
                    MGK_HIP_CHECK(hipEventElapsedTime(&ms, cyc_ev0, cyc_ev1));
                    cyc_ms += ms;
                    cyc_launches += 1;
                    cyc_steps += hctl->inner - inner_before;
                }
                inner_before = hctl->inner;
                if (hctl->done || enqueued >= max_iter) { final_done = true; break; }
                batch = std::min(max_iter - enqueued, std::max(refine ? 1 : 2, batch / 2));
            }
        } else if (warm) {
            // no iteration: the reference still applies the warm-start shift (:181-191)
            ls_enqueue_start(p, LS_START_APPLY, stream);
        }
        if (!final_done) {
            ls_enqueue_final(p, stream);
            MGK_HIP_CHECK(hipMemcpyAsync(hctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
            MGK_HIP_CHECK(hipStreamSynchronize(stream));
        }
        if (event && !hkctl->go) return 0;
        absorb_solve(tol);
        if (diff_out) *diff_out = hctl->diff;
        return hctl->iters;
    }

    // Solver statistics and the field mean after a solve whose control block
    // is in hctl.
    void absorb_solve(double tol) {
        const bool refine = (precision == MGK_FP32_REFINE);
        last_iters = hctl->iters;
        last_inner = hctl->inner;
        predicted = refine ? std::max(1, last_iters) : std::max(2, last_iters);
        for (int k = 0; k < g.C; ++k) mean_eps[k] = hctl->mean_eps[k];
        mean_valid = true;
        if (refine) tune_inner(tol);
    }

    // eps_macro / sig_macro after an event attempt that produced an event:
    // eps_macro advanced by strain_rate * t_wait, then replaced by the field
    // mean (elastic_run's alias, kmc_solver.hpp:334-339).
    void event_macros(const mgk_event_in& in, double t_wait, double* eps_macro, double* sig_macro) {
        const int Dd = D();
        const double x = in.strain_rate * t_wait;
        const int drv = in.drv_i + Dd * in.drv_j;
        double a[9];
        for (int q = 0; q < Dd * Dd; ++q) a[q] = eps_macro[q] + ((q == drv) ? 1.0 : 0.0) * x;
        const int ns = nskew();
        double bar_sym[6], bar_skew[3] = {0, 0, 0};
        split(a, bar_sym, bar_skew);
        for (int w = 0; w < ns; ++w) W[w] = bar_skew[w];
        double sk_sig[3];
        for (int w = 0; w < ns; ++w) sk_sig[w] = 2.0 * mean_mu * bar_skew[w];
        join(hctl->mean_eps, bar_skew, eps_macro);
        join(hctl->mean_sig, sk_sig, sig_macro);
    }

    // CG steps from which a step is speculative: the fewest any of the last
    // few single-cycle solves used (no history: none).  A speculative step
    // tests the control block before loading its inputs, so that past
    // convergence it costs little more than its launch.
    int spec_from() const {
        int mn = 1 << 30;
        for (int h : inner_hist)
            if (h > 0) mn = std::min(mn, h);
        return mn;
    }
    int spec_from2() const {
        int mn = 1 << 30;
        for (int h : inner_hist2)
            if (h > 0) mn = std::min(mn, h);
        return mn;
    }
    static int cg_flags(int j, int spec_from, int J) {
        const int f = j == 0 ? LS_FLAG_CG_FIRST : (j == 1 ? LS_FLAG_CG_UPDATE1 : LS_FLAG_CG_UPDATE);
        return f | (j >= spec_from ? LS_FLAG_SPEC : 0) | ((j & 1) ? LS_FLAG_ODD : 0) | (j == J - 1 ? LS_FLAG_LAST : 0) |
               (j << LS_STEP_SHIFT);
    }

    // CG steps per refinement cycle for the next solve, the first cycle and
    // the later ones apart: one more than the most any of the last few cycles
    // of that kind used (steps past convergence return at once, but each still
    // costs a launch); a cycle that ran out of steps without a decision makes
    // the next ones of its kind longer.
    void tune_inner(double tol) {
        (void)tol;
        const Ctl& h = *hctl;
        int steps[2] = {h.cyc_steps[0], h.cyc_steps[1]};
        const int last = h.inner - h.cycle_start;   // the cycle the solve ended in
        if (last > 0 && h.iters >= 1) {
            const int k = (h.iters <= 1) ? 0 : 1;
            steps[k] = std::max(steps[k], last);
        }
        if (h.cyc_exhausted[0] || (!h.done && h.iters <= 1)) {
            inner_per_outer = std::min(40, inner_per_outer + 4);
        } else if (steps[0] > 0) {
            inner_hist[inner_pos++ % 4] = steps[0];
            int mx = 0;
            for (int v : inner_hist) mx = std::max(mx, v);
            inner_per_outer = std::max(4, std::min(40, mx + 1));
        }
        if (h.cyc_exhausted[1] || (!h.done && h.iters > 1)) {
            inner_per_outer2 = std::min(40, inner_per_outer2 + 2);
        } else if (steps[1] > 0) {
            inner_hist2[inner_pos2++ % 4] = steps[1];
            int mx = 0;
            for (int v : inner_hist2) mx = std::max(mx, v);
            inner_per_outer2 = std::max(2, std::min(40, mx + 1));
        }
    }
};

namespace {

// ---------------------------------------------------------------------------
// Autotuner.  On the first context of a (grid, precision) the block-shape
// variants of the passes that have them (pass A's line pairs, the M passes'
// columns, for FP64 iterations and for five-pass FP32 CG steps on 3D lines of
// 32 points or more) are timed on the new context's buffers, and on grids
// that run the persistent CG cycle a cold solve with it is timed against one
// with the launched plane passes.  A variant replaces the default only when
// it is more than 3% faster.  The choice is kept for the process; no variant
// changes a result (a line's or a column's arithmetic does not depend on the
// block shape, and the cycle equals the plane passes bit for bit).
// MGK_AUTOTUNE=0 keeps the defaults.
// ---------------------------------------------------------------------------
struct TuneRecord {
    LsTune tune{};
    int use_cycle = 1;
    int tuned = 0;             // 1: timed in this process (0: defaults)
    double ms[10] = {0};       // A FP64 v0 v1, B/D FP64 v0 v1, A FP32 v0 v1, B/D FP32 v0 v1, solve cycle, solve plane
};
std::map<std::vector<int>, TuneRecord> g_tunes;
long g_tune_runs = 0;

// Average ms of `reps` launches of one pass variant on the context's stream.
double time_variant(mgk_ctx& c, int pass, int mode, const LsTune& t, int reps) {
    std::memset(c.hctl, 0, sizeof(Ctl));
    c.hctl->iters = 1;
    c.hctl->max_iter = 1 << 30;
    MGK_HIP_CHECK(hipMemcpyAsync(c.ctl, c.hctl, sizeof(Ctl), hipMemcpyHostToDevice, c.stream));
    LsParams q = c.ls_params();
    q.corr = (mode == LS_CG) ? c.corr : nullptr;
    q.r32 = (mode == LS_CG) ? c.cg : nullptr;
    LsLaunch p = c.lsl(1, &q);
    p.tune = t;
    ls_enqueue_pass(p, pass, mode, c.stream);   // warm-up launch
    MGK_HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    for (int r = 0; r < reps; ++r) ls_enqueue_pass(p, pass, mode, c.stream);
    MGK_HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    MGK_HIP_CHECK(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    MGK_HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    return (double)ms / reps;
}

// Wall ms of one cold LS solve (unit uniaxial strain) on the context.
double time_cold_solve(mgk_ctx& c) {
    const double bar[6] = {1e-3, 0, 0, 0, 0, 0};
    MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
    MGK_HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    c.solve(bar, 0.0, false, 200, 1e-8, nullptr);
    MGK_HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    MGK_HIP_CHECK(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    MGK_HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    return (double)ms;
}

void autotune(mgk_ctx& c) {
    const Geom& g = c.g;
    const std::vector<int> key{g.dims, g.nx, g.ny, g.nz, c.precision};
    auto it = g_tunes.find(key);
    if (it == g_tunes.end()) {
        TuneRecord rec;
        const char* e = std::getenv("MGK_AUTOTUNE");
        if (!(e && e[0] == '0')) {
            const bool refine = (c.precision == MGK_FP32_REFINE);
            const bool p64 = ls_plane64_path(g);   // FP64 iterations on the plane passes: no pass A or M pass
            const bool shapes_a = (g.C == 6 && g.nl >= 32);
            const bool shapes_m = (g.C == 6 && g.nm >= 32);
            const bool cg5 = refine && !ls_plane_path(g);   // FP32 CG steps on the five-pass pipeline
            auto pick = [](double d, double v) { return v < 0.97 * d ? 1 : 0; };
            for (int m = 0; m < 2; ++m) {
                if (m == 1 && !cg5) continue;
                if (m == 0 && p64) continue;
                const int mode = (m == 0) ? LS_FULL : LS_CG;
                if (shapes_a) {
                    LsTune t = rec.tune;
                    t.fwd[m] = 0;
                    const double d = time_variant(c, (m == 0) ? 0 : 5, mode, t, 10);
                    t.fwd[m] = 1;
                    const double v = time_variant(c, (m == 0) ? 0 : 5, mode, t, 10);
                    rec.ms[4 * m] = d;
                    rec.ms[4 * m + 1] = v;
                    rec.tune.fwd[m] = pick(d, v);
                }
                if (shapes_m) {
                    LsTune t = rec.tune;
                    t.mid[m] = 0;
                    const double d = time_variant(c, 1, mode, t, 10);
                    t.mid[m] = 1;
                    const double v = time_variant(c, 1, mode, t, 10);
                    rec.ms[4 * m + 2] = d;
                    rec.ms[4 * m + 3] = v;
                    rec.tune.mid[m] = pick(d, v);
                }
            }
            c.tune = rec.tune;
            if (refine && ls_cycle_supported(g)) {
                // each pipeline once for warm-up, then timed
                double best[2] = {0, 0};
                for (int round = 0; round < 2; ++round)
                    for (int uc = 1; uc >= 0; --uc) {
                        c.use_cycle = uc;
                        best[uc] = time_cold_solve(c);
                    }
                rec.ms[8] = best[1];
                rec.ms[9] = best[0];
                rec.use_cycle = (best[0] < 0.97 * best[1]) ? 0 : 1;
            }
            rec.tuned = 1;
            ++g_tune_runs;
        }
        it = g_tunes.emplace(key, rec).first;
    }
    c.tune = it->second.tune;
    c.use_cycle = it->second.use_cycle;
}

// Back to a fresh context after the autotuner's launches: fields, CG
// buffers, control block and the solver statistics.
void reset_after_tuning(mgk_ctx& c) {
    const Geom& g = c.g;
    const long N = g.N;
    const int C = g.C;
    MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
    MGK_HIP_CHECK(hipMemset(c.eps, 0, sizeof(double) * C * N));
    MGK_HIP_CHECK(hipMemset(c.sig, 0, sizeof(double) * C * N));
    MGK_HIP_CHECK(hipMemset(c.spec, 0, sizeof(double) * 2 * C * g.NS));
    MGK_HIP_CHECK(hipMemset(c.cg, 0, sizeof(float) * 4 * C * N));
    MGK_HIP_CHECK(hipMemset(c.corr, 0, sizeof(float) * C * N));
    MGK_HIP_CHECK(hipMemset(c.ctl, 0, sizeof(Ctl)));
    std::memset(c.hctl, 0, sizeof(Ctl));
    c.last_iters = c.last_inner = 0;
    c.predicted = 4;
    c.inner_per_outer = 12;
    c.inner_per_outer2 = 4;
    for (int k = 0; k < 4; ++k) c.inner_hist[k] = c.inner_hist2[k] = 0;
    c.inner_pos = c.inner_pos2 = 0;
    c.mean_valid = false;
    for (int k = 0; k < 6; ++k) c.mean_eps[k] = 0.0;
    c.lsp_up[0] = c.lsp_up[1] = false;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* mgk_last_error(void) { return g_last_error.c_str(); }
int mgk_abi_version(void) { return MGK_ABI_VERSION; }

int mgk_create(int dims, int nx, int ny, int nz, double pixel, int plane_mode, int precision, const double* E,
               const double* nu, mgk_ctx** out) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!out || !E || !nu) throw ArgError("null argument");
        *out = nullptr;
        if (dims != 2 && dims != 3) throw ArgError("dims must be 2 or 3");
        if (dims == 2) nz = 1;
        if (nx < 1 || ny < 1 || nz < 1) throw ArgError("grid too small");
        if (dims == 2 && (nx < 2 || ny < 2)) throw ArgError("grid too small");
        if (dims == 3 && (nx == 1) + (ny == 1) + (nz == 1) > 1) throw ArgError("grid too small: at most one 3D axis of length 1");
        if (precision != MGK_FP64 && precision != MGK_FP32_REFINE) throw ArgError("bad precision");
        const Geom g = make_geom(dims, nx, ny, nz);
        if (g.nl % 2 != 0) throw UnsupportedError("the contiguous axis length must be even");
        for (int n : {g.nf, g.nl})
            if (!ls_length_supported(n)) throw UnsupportedError("FFT length " + std::to_string(n) + " not built");
        if (g.nm > 1 && !ls_length_supported(g.nm)) throw UnsupportedError("FFT length " + std::to_string(g.nm) + " not built");

        mgk_ctx* c = new mgk_ctx();
        try {
            c->g = g;
            c->pixel = pixel;
            c->plane_mode = plane_mode;
            c->precision = precision;
            MGK_HIP_CHECK(hipGetDevice(&c->device));
            MGK_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            MGK_HIP_CHECK(hipEventCreate(&c->ev0));
            MGK_HIP_CHECK(hipEventCreate(&c->ev1));
            MGK_HIP_CHECK(hipEventCreate(&c->cyc_ev0));
            MGK_HIP_CHECK(hipEventCreate(&c->cyc_ev1));
            const long N = g.N;
            // material: mean medium and per-voxel Lame constants
            // (linear_elastic_solver.hpp:163-178, :447-462)
            double Es = 0.0, ns = 0.0;
            for (long i = 0; i < N; ++i) { Es += E[i]; ns += nu[i]; }
            lame(dims, plane_mode, Es / N, ns / N, c->lam0, c->mu0);
            std::vector<double> hl(N), hm(N);
            std::vector<float> hdl(N), hdm(N);
            double msum = 0.0;
            for (long i = 0; i < N; ++i) {
                lame(dims, plane_mode, E[i], nu[i], hl[i], hm[i]);
                hdl[i] = (float)(hl[i] - c->lam0);
                hdm[i] = (float)(hm[i] - c->mu0);
                msum += hm[i];
            }
            c->mean_mu = msum / N;
            c->h_mu = hm;
            c->material_hash = fnv1a(nu, N * sizeof(double), fnv1a(E, N * sizeof(double)));
            c->E = c->dal<double>(N);
            c->nu = c->dal<double>(N);
            c->lam = c->dal<double>(N);
            c->mu = c->dal<double>(N);
            c->dlam32 = c->dal<float>(N);
            c->dmu32 = c->dal<float>(N);
            MGK_HIP_CHECK(hipMemcpy(c->E, E, N * sizeof(double), hipMemcpyHostToDevice));
            MGK_HIP_CHECK(hipMemcpy(c->nu, nu, N * sizeof(double), hipMemcpyHostToDevice));
            MGK_HIP_CHECK(hipMemcpy(c->lam, hl.data(), N * sizeof(double), hipMemcpyHostToDevice));
            MGK_HIP_CHECK(hipMemcpy(c->mu, hm.data(), N * sizeof(double), hipMemcpyHostToDevice));
            MGK_HIP_CHECK(hipMemcpy(c->dlam32, hdl.data(), N * sizeof(float), hipMemcpyHostToDevice));
            MGK_HIP_CHECK(hipMemcpy(c->dmu32, hdm.data(), N * sizeof(float), hipMemcpyHostToDevice));
            const int C = g.C;
            c->eps = c->dal<double>((size_t)C * N);
            c->sig = c->dal<double>((size_t)C * N);
            c->epsp = c->dal<double>((size_t)C * N);
            MGK_HIP_CHECK(hipMemset(c->eps, 0, sizeof(double) * C * N));
            MGK_HIP_CHECK(hipMemset(c->sig, 0, sizeof(double) * C * N));
            MGK_HIP_CHECK(hipMemset(c->epsp, 0, sizeof(double) * C * N));
            c->spec = c->dal<double>((size_t)2 * C * g.NS);
            c->cg = c->dal<float>((size_t)4 * C * N);
            c->corr = c->dal<float>((size_t)C * N);
            c->partials = c->dal<double>((size_t)ls_partial_slots(g) + 64);
            // the LS and KMC control blocks side by side, so one copy brings both back
            CtlPair* pair = c->dal<CtlPair>(1);
            CtlPair* hpair = halloc<CtlPair>(1);
            MGK_HIP_CHECK(hipMemset(pair, 0, sizeof(CtlPair)));
            std::memset(hpair, 0, sizeof(CtlPair));
            c->ctl = &pair->ls;
            c->hctl = &hpair->ls;
            c->kctl = &pair->kmc;
            c->hkctl = &hpair->kmc;
            c->d_lsp = c->dal<LsParams>(2);
            c->h_lsp = halloc<LsParams>(2);
            c->d_kp = c->dal<KmcParams>(1);
            c->h_kp = halloc<KmcParams>(1);
            c->d_ea = c->dal<EventArgs>(1);
            c->h_ea = halloc<EventArgs>(1);
            MGK_HIP_CHECK(hipMemset(c->ctl, 0, sizeof(Ctl)));
            const int lens[3] = {g.nf, g.nm, g.nl};
            for (int a = 0; a < 3; ++a) {
                auto t64 = twiddle_table<double>(lens[a]);
                auto t32 = twiddle_table<float>(lens[a]);
                c->tw64[a] = c->dal<double>(t64.size());
                c->tw32[a] = c->dal<float>(t32.size());
                MGK_HIP_CHECK(hipMemcpy(c->tw64[a], t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
                MGK_HIP_CHECK(hipMemcpy(c->tw32[a], t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice));
            }
            MGK_HIP_CHECK(hipDeviceSynchronize());
            const long runs = g_tune_runs;
            autotune(*c);
            if (g_tune_runs != runs) reset_after_tuning(*c);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
        return MGK_OK;
    });
}

int mgk_destroy(mgk_ctx* ctx) {
    if (!ctx) return MGK_OK;
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
        return MGK_OK;
    });
}

int mgk_set_precision(mgk_ctx* ctx, int precision) {
    return guarded([&]() -> int {
        if (!ctx) throw ArgError("null context");
        if (precision != MGK_FP64 && precision != MGK_FP32_REFINE) throw ArgError("bad precision");
        ctx->precision = precision;
        return MGK_OK;
    });
}

int mgk_ls_solve(mgk_ctx* ctx, const double* eps_bar, double* eps, int warm_start, const double* eps_plastic,
                 double* sig_out, double* eps_macro, double* sig_macro, int max_iter, double tol, int* iters_out,
                 double* diff_out) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !eps_bar || !eps || !eps_macro || !sig_macro) throw ArgError("null argument");
        if (max_iter < 0) throw ArgError("max_iter < 0");
        mgk_ctx& c = *ctx;
        const long N = c.g.N;
        const int ns = c.nskew();
        double bar_sym[6], bar_skew[3] = {0, 0, 0};
        c.split(eps_bar, bar_sym, bar_skew);   // copy first: eps_macro may alias eps_bar
        // plastic strain: symmetric part to the device, antisymmetric part kept here
        std::vector<double> pskew;
        if (eps_plastic) {
            c.upload_tensor(eps_plastic, c.epsp, &pskew);
            c.epsp_set = true;
        } else {
            c.epsp_set = false;
        }
        std::vector<double> eskew;
        // ||eps||^2 of the uniform antisymmetric part (each w appears at (i,j) and (j,i))
        double norm_extra = 0.0;
        for (int w = 0; w < ns; ++w) norm_extra += 2.0 * (double)N * bar_skew[w] * bar_skew[w];
        if (warm_start) {
            c.upload_tensor(eps, c.eps, &eskew);
            // The device keeps symmetric parts; the antisymmetric part after the
            // warm-start shift is W_v + W_bar - mean(W), uniform only if W_v is.
            for (long v = 1; v < N; ++v)
                for (int w = 0; w < ns; ++w)
                    if (eskew[(size_t)v * ns + w] != eskew[w])
                        throw ArgError("the antisymmetric part of a warm-start strain field must be uniform");
        }
        if (warm_start) c.mean_valid = false;   // a new field was uploaded
        int it = c.solve(bar_sym, norm_extra, warm_start != 0, max_iter, tol, diff_out);
        if (iters_out) *iters_out = it;
        for (int w = 0; w < ns; ++w) c.W[w] = bar_skew[w];
        // outputs
        c.download_tensor(c.eps, eps, [&](long, double* sk) { for (int w = 0; w < ns; ++w) sk[w] = bar_skew[w]; });
        const std::vector<double>& hm = c.h_mu;
        if (sig_out) {
            c.download_tensor(c.sig, sig_out, [&](long v, double* sk) {
                for (int w = 0; w < ns; ++w) {
                    const double pw = pskew.empty() ? 0.0 : pskew[(size_t)v * ns + w];
                    sk[w] = 2.0 * hm[v] * (bar_skew[w] - pw);
                }
            });
        }
        double sk_sig[3] = {0, 0, 0};
        for (int w = 0; w < ns; ++w) {
            double s = 0.0;
            for (long v = 0; v < N; ++v) {
                const double pw = pskew.empty() ? 0.0 : pskew[(size_t)v * ns + w];
                s += 2.0 * hm[v] * (bar_skew[w] - pw);
            }
            sk_sig[w] = s / N;
        }
        c.join(c.hctl->mean_eps, bar_skew, eps_macro);
        c.join(c.hctl->mean_sig, sk_sig, sig_macro);
        return MGK_OK;
    });
}

static long field_len(const mgk_ctx& c, int which) {
    const long N = c.g.N;
    const int F = c.full(), M = c.kp.M;
    switch (which) {
        case MGK_FIELD_EPS: case MGK_FIELD_SIG: case MGK_FIELD_EPS_PLASTIC: case MGK_FIELD_PREV_DIR: return N * F;
        case MGK_FIELD_Q: case MGK_FIELD_Q0: return N * M;
        case MGK_FIELD_CATALOG: return N * M * F;
        case MGK_FIELD_SOFT: return 2 * N;
        case MGK_FIELD_LAST_EVENT_TIME: case MGK_FIELD_E: case MGK_FIELD_NU: return N;
    }
    return -1;
}

static bool kmc_field(int which) {
    return which == MGK_FIELD_Q || which == MGK_FIELD_Q0 || which == MGK_FIELD_CATALOG || which == MGK_FIELD_SOFT ||
           which == MGK_FIELD_LAST_EVENT_TIME || which == MGK_FIELD_PREV_DIR;
}

int mgk_set_field(mgk_ctx* ctx, int which, const double* host) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !host) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (field_len(c, which) < 0) throw ArgError("unknown field");
        if (kmc_field(which) && !c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        const long N = c.g.N;
        const int ns = c.nskew();
        std::vector<double> sk;
        switch (which) {
            case MGK_FIELD_EPS: {
                c.mean_valid = false;
                c.upload_tensor(host, c.eps, &sk);
                for (int w = 0; w < ns; ++w) c.W[w] = sk.empty() ? 0.0 : sk[w];
                for (long v = 0; v < N; ++v)
                    for (int w = 0; w < ns; ++w)
                        if (sk[(size_t)v * ns + w] != c.W[w]) throw ArgError("antisymmetric part of eps must be uniform");
                break;
            }
            case MGK_FIELD_SIG: c.upload_tensor(host, c.sig, nullptr); break;
            case MGK_FIELD_EPS_PLASTIC: {
                c.upload_tensor(host, c.epsp, &sk);
                for (double s : sk)
                    if (s != 0.0) throw ArgError("eps_plastic must be symmetric on the device path");
                c.epsp_set = true;
                break;
            }
            case MGK_FIELD_PREV_DIR: c.upload_tensor(host, c.K.prev_dir, nullptr); break;
            case MGK_FIELD_CATALOG: {
                // N*M tensors -> [C][N*M]; when every voxel holds voxel 0's M
                // strains, one shared [C][M] copy instead
                const long nm = N * c.kp.M;
                const int C = c.g.C, F = c.full(), M = c.kp.M;
                bool shared = true;
                for (long i = M; i < nm && shared; ++i)
                    for (int k = 0; k < F; ++k)
                        if (host[i * F + k] != host[(i % M) * F + k]) { shared = false; break; }
                const long n = shared ? M : nm;
                std::vector<double> h((size_t)n * C);
                double sym[6], skw[3];
                for (long i = 0; i < n; ++i) {
                    c.split(host + i * F, sym, skw);
                    for (int k = 0; k < C; ++k) h[(size_t)k * n + i] = sym[k];
                }
                c.ensure_cat(h.size());
                MGK_HIP_CHECK(hipMemcpy(c.K.cat, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
                c.K.cat_shared = shared ? 1 : 0;
                if (shared) c.shared_cat = h;
                else c.shared_cat.clear();
                break;
            }
            case MGK_FIELD_Q: MGK_HIP_CHECK(hipMemcpy(c.K.Q, host, N * c.kp.M * sizeof(double), hipMemcpyHostToDevice)); break;
            case MGK_FIELD_Q0: MGK_HIP_CHECK(hipMemcpy(c.K.Q0, host, N * c.kp.M * sizeof(double), hipMemcpyHostToDevice)); break;
            case MGK_FIELD_SOFT: {
                std::vector<double> h(2 * N);
                for (long v = 0; v < N; ++v) { h[v] = host[2 * v]; h[N + v] = host[2 * v + 1]; }
                MGK_HIP_CHECK(hipMemcpy(c.K.soft, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
                break;
            }
            case MGK_FIELD_LAST_EVENT_TIME:
                MGK_HIP_CHECK(hipMemcpy(c.K.last_t, host, N * sizeof(double), hipMemcpyHostToDevice));
                break;
            default: throw ArgError("field is read-only");
        }
        MGK_HIP_CHECK(hipDeviceSynchronize());
        return MGK_OK;
    });
}

int mgk_get_field(mgk_ctx* ctx, int which, double* host) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !host) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (field_len(c, which) < 0) throw ArgError("unknown field");
        if (kmc_field(which) && !c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        const long N = c.g.N;
        const int ns = c.nskew();
        const double zero[3] = {0, 0, 0};
        switch (which) {
            case MGK_FIELD_EPS:
                c.download_tensor(c.eps, host, [&](long, double* sk) { for (int w = 0; w < ns; ++w) sk[w] = c.W[w]; });
                break;
            case MGK_FIELD_SIG:
                c.download_tensor(c.sig, host, [&](long v, double* sk) {
                    for (int w = 0; w < ns; ++w) sk[w] = 2.0 * c.h_mu[v] * c.W[w];
                });
                break;
            case MGK_FIELD_EPS_PLASTIC:
                c.download_tensor(c.epsp, host, [&](long, double* sk) { for (int w = 0; w < ns; ++w) sk[w] = zero[w]; });
                break;
            case MGK_FIELD_PREV_DIR:
                c.download_tensor(c.K.prev_dir, host, [&](long, double* sk) { for (int w = 0; w < ns; ++w) sk[w] = zero[w]; });
                break;
            case MGK_FIELD_CATALOG: {
                const long nm = N * c.kp.M;
                const int C = c.g.C, F = c.full(), M = c.kp.M;
                const long n = c.K.cat_shared ? M : nm;
                std::vector<double> h((size_t)n * C);
                MGK_HIP_CHECK(hipMemcpy(h.data(), c.K.cat, h.size() * sizeof(double), hipMemcpyDeviceToHost));
                double sym[6];
                for (long i = 0; i < nm; ++i) {
                    const long src = c.K.cat_shared ? i % M : i;
                    for (int k = 0; k < C; ++k) sym[k] = h[(size_t)k * n + src];
                    c.join(sym, zero, host + i * F);
                }
                break;
            }
            case MGK_FIELD_Q: MGK_HIP_CHECK(hipMemcpy(host, c.K.Q, N * c.kp.M * sizeof(double), hipMemcpyDeviceToHost)); break;
            case MGK_FIELD_Q0: MGK_HIP_CHECK(hipMemcpy(host, c.K.Q0, N * c.kp.M * sizeof(double), hipMemcpyDeviceToHost)); break;
            case MGK_FIELD_SOFT: {
                std::vector<double> h(2 * N);
                MGK_HIP_CHECK(hipMemcpy(h.data(), c.K.soft, h.size() * sizeof(double), hipMemcpyDeviceToHost));
                for (long v = 0; v < N; ++v) { host[2 * v] = h[v]; host[2 * v + 1] = h[N + v]; }
                break;
            }
            case MGK_FIELD_LAST_EVENT_TIME:
                MGK_HIP_CHECK(hipMemcpy(host, c.K.last_t, N * sizeof(double), hipMemcpyDeviceToHost));
                break;
            case MGK_FIELD_E: MGK_HIP_CHECK(hipMemcpy(host, c.E, N * sizeof(double), hipMemcpyDeviceToHost)); break;
            case MGK_FIELD_NU: MGK_HIP_CHECK(hipMemcpy(host, c.nu, N * sizeof(double), hipMemcpyDeviceToHost)); break;
        }
        return MGK_OK;
    });
}

int mgk_kmc_init(mgk_ctx* ctx, const mgk_kmc_params* kp) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !kp) throw ArgError("null argument");
        if (kp->M < 1) throw ArgError("M must be >= 1");
        mgk_ctx& c = *ctx;
        const long N = c.g.N;
        const long nm = N * kp->M;
        if (nm > (long)std::numeric_limits<int>::max()) throw ArgError("N*M exceeds the int index range of the reference");
        if (c.kmc) {
            auto f = [](void* p) { if (p) (void)hipFree(p); };
            f(c.K.Q); f(c.K.Q0); f(c.K.cat); f(c.K.soft); f(c.K.last_t); f(c.K.prev_dir); f(c.K.bsum);
            f(c.K.bincl); f(c.K.bunst); f(c.K.bunst_off); f(c.K.binc); f(c.kstage); f(c.unst_idx); f(c.unst_q);
            f(c.fstage); f(c.fidx);
            if (c.hkstage) (void)hipHostFree(c.hkstage);
            if (c.hfstage) (void)hipHostFree(c.hfstage);
            if (c.hfidx) (void)hipHostFree(c.hfidx);
            c.dev_bytes -= c.kmc_bytes;
            c.kmc_bytes = 0;
            c.cat_cap = 0;
        }
        c.kp = *kp;
        const int C = c.g.C;
        KmcParams& K = c.K;
        K = KmcParams{};
        K.g = c.g;
        KmcPhys& ph = K.ph;
        ph.M = kp->M;
        ph.directional = kp->softening_directional;
        ph.instability_kmc = kp->instability_kmc;
        ph.cap = kp->softening_cap;
        ph.jp = kp->jp;
        ph.jt = kp->jt;
        ph.nfrac = kp->neighbor_softening_fraction;
        ph.tau = kp->tau;
        ph.tau_inf = std::isinf(kp->tau) ? 1 : 0;
        ph.volume = kp->volume;
        ph.eff_volume = kp->scale_rate_by_volume ? kp->volume : 1.0;   // kmc_solver.hpp:784
        ph.nu0 = kp->nu0;
        const double kB = 8.617e-5;                                       // kmc_solver.hpp:785
        ph.beta = (kp->temperature > 0.0) ? (1.0 / (kB * kp->temperature)) : std::numeric_limits<double>::infinity();
        ph.thr = kp->stability_threshold;
        K.sig = c.sig;
        K.eps_p = c.epsp;
        K.Q = c.dal<double>(nm, true);
        K.Q0 = c.dal<double>(nm, true);
        K.soft = c.dal<double>(2 * N, true);
        K.last_t = c.dal<double>(N, true);
        K.prev_dir = c.dal<double>((size_t)C * N, true);
        const int nb = kmc_nblocks(nm);
        K.bsum = c.dal<double>(nb, true);
        K.bincl = c.dal<double>(nb, true);
        K.bunst = c.dal<int>(nb, true);
        K.bunst_off = c.dal<int>(nb, true);
        K.binc = c.dal<int>(nb, true);
        K.kctl = c.kctl;   // allocated with the LS control block
        c.kstage = c.dal<double>((size_t)kp->M * (C + 1), true);
        c.hkstage = halloc<double>((size_t)kp->M * (C + 1));
        K.staging = c.kstage;
        c.fstage = c.dal<double>((size_t)kFlipChunk * kp->M * (C + 1), true);
        c.hfstage = halloc<double>((size_t)kFlipChunk * kp->M * (C + 1));
        c.fidx = c.dal<int>(kFlipChunk, true);
        c.hfidx = halloc<int>(kFlipChunk);
        c.unst_idx = c.dal<int>(nm, true);
        c.unst_q = c.dal<double>(nm, true);
        K.unst_idx = c.unst_idx;
        K.unst_q = c.unst_q;
        MGK_HIP_CHECK(hipMemset(K.Q, 0, nm * sizeof(double)));
        MGK_HIP_CHECK(hipMemset(K.Q0, 0, nm * sizeof(double)));
        // the zero catalog every voxel starts with is one shared copy
        c.ensure_cat((size_t)C * kp->M);
        MGK_HIP_CHECK(hipMemset(K.cat, 0, (size_t)C * kp->M * sizeof(double)));
        K.cat_shared = 1;
        c.shared_cat.assign((size_t)C * kp->M, 0.0);
        MGK_HIP_CHECK(hipMemset(K.soft, 0, 2 * N * sizeof(double)));
        MGK_HIP_CHECK(hipMemset(K.prev_dir, 0, (size_t)C * N * sizeof(double)));
        MGK_HIP_CHECK(hipMemset(c.kctl, 0, sizeof(KmcCtl)));
        std::vector<double> ninf(N, -std::numeric_limits<double>::infinity());
        MGK_HIP_CHECK(hipMemcpy(K.last_t, ninf.data(), N * sizeof(double), hipMemcpyHostToDevice));
        MGK_HIP_CHECK(hipMemset(c.epsp, 0, sizeof(double) * C * N));
        c.epsp_set = true;
        MGK_HIP_CHECK(hipDeviceSynchronize());
        c.kmc = true;
        return MGK_OK;
    });
}

int mgk_elastic_run(mgk_ctx* ctx, double* eps_macro, double* sig_macro, int max_iter, double tol, int* iters_out,
                    double* diff_out) {
    return guarded([&]() -> int {
        if (!ctx || !eps_macro || !sig_macro) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        const int ns = c.nskew();
        double bar_sym[6], bar_skew[3] = {0, 0, 0};
        c.split(eps_macro, bar_sym, bar_skew);
        double norm_extra = 0.0;
        for (int w = 0; w < ns; ++w) norm_extra += 2.0 * (double)c.g.N * bar_skew[w] * bar_skew[w];
        // warm start always: the KMC fields are sized N (kmc_solver.hpp:197)
        int it = c.solve(bar_sym, norm_extra, true, max_iter, tol, diff_out);
        if (iters_out) *iters_out = it;
        for (int w = 0; w < ns; ++w) c.W[w] = bar_skew[w];
        double sk_sig[3];
        for (int w = 0; w < ns; ++w) sk_sig[w] = 2.0 * c.mean_mu * bar_skew[w];
        c.join(c.hctl->mean_eps, bar_skew, eps_macro);
        c.join(c.hctl->mean_sig, sk_sig, sig_macro);
        return MGK_OK;
    });
}

int mgk_update_rates(mgk_ctx* ctx, double time, double* total_rate, int* n_unstable, int* n_included) {
    return guarded([&]() -> int {
        if (!ctx) throw ArgError("null context");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        kmc_enqueue_rates(c.kml(), time, nullptr, c.stream);
        MGK_HIP_CHECK(hipMemcpyAsync(c.hkctl, c.kctl, sizeof(KmcCtl), hipMemcpyDeviceToHost, c.stream));
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        if (total_rate) *total_rate = c.hkctl->total_rate;
        if (n_unstable) *n_unstable = c.hkctl->n_unstable;
        if (n_included) *n_included = c.hkctl->n_included;
        return MGK_OK;
    });
}

int mgk_select(mgk_ctx* ctx, double u2, int* flat_idx, double* q_sel) {
    return guarded([&]() -> int {
        if (!ctx || !flat_idx) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        kmc_enqueue_select(c.kml(), u2, nullptr, c.stream);
        MGK_HIP_CHECK(hipMemcpyAsync(c.hkctl, c.kctl, sizeof(KmcCtl), hipMemcpyDeviceToHost, c.stream));
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        if (c.hkctl->ambiguous) {
            kmc_enqueue_select_fallback(c.kml(), 0, c.stream);
            MGK_HIP_CHECK(hipMemcpyAsync(c.hkctl, c.kctl, sizeof(KmcCtl), hipMemcpyDeviceToHost, c.stream));
            MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        }
        *flat_idx = c.hkctl->sel;
        if (q_sel) *q_sel = c.hkctl->q_sel;
        return MGK_OK;
    });
}

int mgk_select_exact(mgk_ctx* ctx, double u2, int* flat_idx, double* q_sel) {
    return guarded([&]() -> int {
        if (!ctx || !flat_idx) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        kmc_enqueue_select(c.kml(), u2, nullptr, c.stream);   // sets u2 and the go flag
        kmc_enqueue_select_fallback(c.kml(), 0, c.stream);
        MGK_HIP_CHECK(hipMemcpyAsync(c.hkctl, c.kctl, sizeof(KmcCtl), hipMemcpyDeviceToHost, c.stream));
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        *flat_idx = c.hkctl->sel;
        if (q_sel) *q_sel = c.hkctl->q_sel;
        return MGK_OK;
    });
}

int mgk_apply_flip(mgk_ctx* ctx, int flat_idx, double time, const double* new_catalog, const double* new_Q0,
                   int first_mode, int n_modes) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx) throw ArgError("null context");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        const long nm = c.g.N * c.kp.M;
        if (flat_idx < 0 || flat_idx >= nm) throw ArgError("flat index out of range");
        if ((new_catalog || new_Q0) && (first_mode < 0 || n_modes < 0 || first_mode + n_modes > c.kp.M))
            throw ArgError("redraw mode range out of range");
        if (!new_catalog && !new_Q0) n_modes = 0;
        const int C = c.g.C, F = c.full(), M = c.kp.M;
        double* h = c.hkstage;
        double skw[3];
        if (new_catalog)
            for (int j = 0; j < n_modes; ++j) c.split(new_catalog + (size_t)j * F, h + (size_t)j * C, skw);
        if (new_catalog && c.K.cat_shared) {
            // a redraw that reproduces the shared strains keeps the shared copy;
            // any other redraw turns the catalog into a per-(voxel, mode) array
            const std::vector<double>& shared = c.shared_cat;   // [C][M], kept since mgk_set_field
            bool same = true;
            for (int j = 0; j < n_modes && same; ++j)
                for (int k = 0; k < C; ++k)
                    if (h[(size_t)j * C + k] != shared[(size_t)k * M + first_mode + j]) { same = false; break; }
            if (same) {
                new_catalog = nullptr;
            } else {
                std::vector<double> full((size_t)C * nm);
                for (int k = 0; k < C; ++k)
                    for (long i = 0; i < nm; ++i) full[(size_t)k * nm + i] = shared[(size_t)k * M + i % M];
                MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
                c.ensure_cat(full.size());
                MGK_HIP_CHECK(hipMemcpyAsync(c.K.cat, full.data(), full.size() * sizeof(double),
                                             hipMemcpyHostToDevice, c.stream));
                MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
                c.K.cat_shared = 0;
            }
        }
        if (!new_catalog && !new_Q0) n_modes = 0;
        if (new_Q0)
            for (int j = 0; j < n_modes; ++j) h[(size_t)n_modes * C + j] = new_Q0[j];
        if (n_modes > 0)
            MGK_HIP_CHECK(hipMemcpyAsync(c.kstage, h, sizeof(double) * n_modes * (C + 1), hipMemcpyHostToDevice, c.stream));
        kmc_enqueue_flip(c.kml(), flat_idx, time, first_mode, n_modes, new_catalog ? 1 : 0, new_Q0 ? 1 : 0, c.stream);
        // the pinned staging buffer is reused by the next call: wait for the copy
        if (n_modes > 0) MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        return MGK_OK;
    });
}

// Redraw values of a selected voxel into the pinned staging buffer, checked
// against the mode range and the shared-catalog mode.  Returns the number of
// modes staged; *has_cat / *has_q0 say which halves the flip applies.
int mgk_apply_flips(mgk_ctx* ctx, int n, const int* flat_idx, double time, const double* new_catalog,
                    const double* new_Q0, int all_modes) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || (n > 0 && !flat_idx)) throw ArgError("null argument");
        if (n < 0) throw ArgError("negative flip count");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        const long nm = c.g.N * c.kp.M;
        const int C = c.g.C, F = c.full(), M = c.kp.M;
        const int nmo = all_modes ? M : 1;
        for (int i = 0; i < n; ++i)
            if (flat_idx[i] < 0 || flat_idx[i] >= nm) throw ArgError("flat index out of range");
        if (n == 0) return MGK_OK;
        double skw[3];
        if (new_catalog && c.K.cat_shared) {
            // as mgk_apply_flip: redraws that all reproduce the shared strains
            // keep the shared copy, any other makes the catalog per-(voxel, mode)
            const std::vector<double>& shared = c.shared_cat;
            bool same = true;
            double h[6];
            for (int i = 0; i < n && same; ++i) {
                const int first = all_modes ? 0 : flat_idx[i] % M;
                for (int j = 0; j < nmo && same; ++j) {
                    c.split(new_catalog + ((size_t)i * nmo + j) * F, h, skw);
                    for (int k = 0; k < C; ++k)
                        if (h[k] != shared[(size_t)k * M + first + j]) { same = false; break; }
                }
            }
            if (same) {
                new_catalog = nullptr;
            } else {
                std::vector<double> full((size_t)C * nm);
                for (int k = 0; k < C; ++k)
                    for (long i = 0; i < nm; ++i) full[(size_t)k * nm + i] = shared[(size_t)k * M + i % M];
                MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
                c.ensure_cat(full.size());
                MGK_HIP_CHECK(hipMemcpyAsync(c.K.cat, full.data(), full.size() * sizeof(double),
                                             hipMemcpyHostToDevice, c.stream));
                MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
                c.K.cat_shared = 0;
            }
        }
        const int has_cat = new_catalog ? 1 : 0, has_q0 = new_Q0 ? 1 : 0;
        const size_t per = (size_t)nmo * (C + 1);
        const KmcLaunch KL = c.kml();
        for (int i0 = 0; i0 < n; i0 += kFlipChunk) {
            const int cnt = std::min(kFlipChunk, n - i0);
            MGK_HIP_CHECK(hipStreamSynchronize(c.stream));   // the pinned buffers may still feed the last chunk
            for (int i = 0; i < cnt; ++i) {
                c.hfidx[i] = flat_idx[i0 + i];
                double* h = c.hfstage + (size_t)i * per;
                for (int j = 0; j < nmo; ++j) {
                    if (has_cat) c.split(new_catalog + ((size_t)(i0 + i) * nmo + j) * F, h + (size_t)j * C, skw);
                    if (has_q0) h[(size_t)nmo * C + j] = new_Q0[(size_t)(i0 + i) * nmo + j];
                }
            }
            MGK_HIP_CHECK(hipMemcpyAsync(c.fidx, c.hfidx, sizeof(int) * cnt, hipMemcpyHostToDevice, c.stream));
            if (has_cat || has_q0)
                MGK_HIP_CHECK(hipMemcpyAsync(c.fstage, c.hfstage, sizeof(double) * per * cnt, hipMemcpyHostToDevice,
                                             c.stream));
            kmc_enqueue_flip_list(KL, c.fidx, c.fstage, cnt, time, all_modes ? 1 : 0, has_cat, has_q0, c.stream);
        }
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        return MGK_OK;
    });
}

static int stage_redraw(mgk_ctx& c, const double* new_catalog, const double* new_Q0, int first_mode, int n_modes,
                        int* has_cat, int* has_q0, double* dest = nullptr) {
    const int C = c.g.C, F = c.full(), M = c.kp.M;
    if ((new_catalog || new_Q0) &&
        (n_modes < 0 || (first_mode >= 0 && first_mode + n_modes > M) || (first_mode < 0 && n_modes != 1)))
        throw ArgError("redraw mode range out of range");
    if (new_catalog && c.K.cat_shared) {
        // the voxel is chosen on the device: only a redraw that reproduces the
        // shared strains (the fixed catalog) keeps the shared copy
        bool same = first_mode >= 0;
        double sym[6], skw[3];
        for (int j = 0; j < n_modes && same; ++j) {
            c.split(new_catalog + (size_t)j * F, sym, skw);
            for (int k = 0; k < C; ++k)
                if (sym[k] != c.shared_cat[(size_t)k * M + first_mode + j]) { same = false; break; }
        }
        if (!same) throw ArgError("a redraw that differs from the shared catalog needs mgk_select + mgk_apply_flip");
        new_catalog = nullptr;
    }
    if (!new_catalog && !new_Q0) n_modes = 0;
    // the pinned staging buffer may still feed the previous call's copy
    if (!dest && c.kctl_pending) MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
    double* h = dest ? dest : c.hkstage;
    double skw[3];
    if (new_catalog)
        for (int j = 0; j < n_modes; ++j) c.split(new_catalog + (size_t)j * F, h + (size_t)j * C, skw);
    if (new_Q0)
        for (int j = 0; j < n_modes; ++j) h[(size_t)n_modes * C + j] = new_Q0[j];
    *has_cat = new_catalog ? 1 : 0;
    *has_q0 = new_Q0 ? 1 : 0;
    return n_modes;
}

int mgk_select_flip(mgk_ctx* ctx, double u2, double time, const double* new_catalog, const double* new_Q0,
                    int first_mode, int n_modes) {
    return guarded([&]() -> int {
        if (!ctx) throw ArgError("null context");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        int has_cat = 0, has_q0 = 0;
        n_modes = stage_redraw(c, new_catalog, new_Q0, first_mode, n_modes, &has_cat, &has_q0);
        kmc_enqueue_select(c.kml(), u2, nullptr, c.stream);
        kmc_enqueue_select_fallback(c.kml(), 1, c.stream);
        if (n_modes > 0)
            MGK_HIP_CHECK(hipMemcpyAsync(c.kstage, c.hkstage, sizeof(double) * n_modes * (c.g.C + 1),
                                         hipMemcpyHostToDevice, c.stream));
        kmc_enqueue_flip(c.kml(), -1, time, first_mode, n_modes, has_cat, has_q0, c.stream);
        c.kctl_pending = true;   // the selection is read back with the next solve's sync
        return MGK_OK;
    });
}

int mgk_kmc_event(mgk_ctx* ctx, const mgk_event_in* in, double* eps_macro, double* sig_macro, mgk_event_out* out) {
    return guarded([&]() -> int {
        if (!ctx || !in || !eps_macro || !sig_macro || !out) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        const int Dd = c.D();
        if (in->drv_i < 0 || in->drv_i >= Dd || in->drv_j < 0 || in->drv_j >= Dd) throw ArgError("driven component out of range");
        if (in->max_iter < 1) throw ArgError("max_iter must be >= 1");
        const int pm = in->patch_mode;
        if (pm < 0 || pm > 2) throw ArgError("patch_mode must be 0, 1 or 2");
        if (pm && !c.patch_ready) throw ArgError("fast patching not initialised (mgk_patch_init)");
        if (pm && !c.unit_ready) throw ArgError("sigma_macro_unit not set (mgk_patch_unit)");
        int has_cat = 0, has_q0 = 0;
        const int n_modes = stage_redraw(c, in->new_catalog, in->new_Q0, in->first_mode, in->n_modes, &has_cat, &has_q0);
        // LS control block: reset on the device by the event decision (init_ctl),
        // which also fills the imposed strain and the shift
        const int start = c.start_mode(true);
        EventArgs ea{};
        ea.time = in->time;
        ea.neg_log_u1 = in->neg_log_u1;
        ea.remaining = in->remaining;
        ea.strain_rate = in->strain_rate;
        for (int k = 0; k < Dd * Dd; ++k) ea.a[k] = eps_macro[k];
        ea.n_vox = (double)c.g.N;
        ea.drv = in->drv_i + Dd * in->drv_j;
        ea.cascade_check = in->cascade_check ? 1 : 0;
        ea.keep_shift = (start == LS_START_KEEP) ? 1 : 0;
        ea.u2 = in->u2;
        ea.active = 1;
        ea.init_ctl = 1;
        ea.max_iter = in->max_iter;
        ea.tol = in->tol;
        for (int k = 0; k < 6; ++k) ea.mean_eps[k] = c.mean_eps[k];
        const KmcLaunch KL = c.kml();
        const LsLaunch L = c.lsl();
        // redraw values: kernel arguments when they fit, else the staging copy
        const bool by_args = n_modes > 0 && kmc_redraw_fits_args(n_modes, c.g.C);
        if (n_modes > 0 && !by_args)
            MGK_HIP_CHECK(hipMemcpyAsync(c.kstage, c.hkstage, sizeof(double) * n_modes * (c.g.C + 1),
                                         hipMemcpyHostToDevice, c.stream));
        // rates, then one block for scan, decision, selection and flip; FP64
        // runs the exact selection pass before its flip
        const bool exact = (c.precision == MGK_FP64);
        kmc_enqueue_rates(KL, in->time, nullptr, c.stream, false);
        kmc_enqueue_event_chain(KL, c.ctl, nullptr, ea, nullptr, !exact, in->first_mode, n_modes, has_cat, has_q0,
                                by_args ? c.hkstage : nullptr, c.stream);
        if (exact) {
            kmc_enqueue_select_fallback(KL, 1, c.stream);
            kmc_enqueue_flip(KL, -1, in->time, in->first_mode, n_modes, has_cat, has_q0, c.stream, 1,
                             by_args ? c.hkstage : nullptr);
        }
        if (pm) kmc_enqueue_patch(KL, c.pp, -1, nullptr, 0.0, 1, in->strain_rate, c.stream);
        if (pm == 1) {
            // patched, not solved: the control block alone comes back; eps_macro
            // advances by strain_rate * t_wait and sig_macro stays (:811-814)
            MGK_HIP_CHECK(hipMemcpyAsync(c.hkctl, c.kctl, sizeof(KmcCtl), hipMemcpyDeviceToHost, c.stream));
            MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
            c.kctl_pending = false;
            const KmcCtl& k = *c.hkctl;
            std::memset(out, 0, sizeof(*out));
            out->outcome = k.outcome;
            out->total_rate = k.total_rate;
            out->n_unstable = k.n_unstable;
            out->n_included = k.n_included;
            out->t_wait = k.t_wait;
            out->flat_idx = k.go ? k.sel : -1;
            out->q_sel = k.go ? k.q_sel : 0.0;
            out->exact_selection = (k.go && k.ambiguous && c.precision == MGK_FP64) ? 1 : 0;
            if (k.outcome == MGK_EVENT_DONE) {
                const double x = in->strain_rate * k.t_wait;
                const int drv = in->drv_i + Dd * in->drv_j;
                for (int q = 0; q < Dd * Dd; ++q) eps_macro[q] += ((q == drv) ? 1.0 : 0.0) * x;
            }
            return MGK_OK;
        }
        ls_enqueue_start(L, start, c.stream);
        c.kctl_pending = true;
        double diff = 0.0;
        const int it = c.run_batches(true, in->max_iter, in->tol, &diff, true);
        const KmcCtl& k = *c.hkctl;
        std::memset(out, 0, sizeof(*out));
        out->outcome = k.outcome;
        out->total_rate = k.total_rate;
        out->n_unstable = k.n_unstable;
        out->n_included = k.n_included;
        out->t_wait = k.t_wait;
        out->flat_idx = k.go ? k.sel : -1;
        out->q_sel = k.go ? k.q_sel : 0.0;
        out->exact_selection = (k.go && k.ambiguous && c.precision == MGK_FP64) ? 1 : 0;
        if (k.outcome != MGK_EVENT_DONE) return MGK_OK;
        out->iters = it;
        out->diff = diff;
        c.event_macros(*in, k.t_wait, eps_macro, sig_macro);
        return MGK_OK;
    });
}

int mgk_last_selection(mgk_ctx* ctx, int* flat_idx, double* q_sel) {
    return guarded([&]() -> int {
        if (!ctx || !flat_idx) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (c.kctl_pending) {
            MGK_HIP_CHECK(hipMemcpyAsync(c.hkctl, c.kctl, sizeof(KmcCtl), hipMemcpyDeviceToHost, c.stream));
            MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
            c.kctl_pending = false;
        }
        *flat_idx = c.hkctl->sel;
        if (q_sel) *q_sel = c.hkctl->q_sel;
        return MGK_OK;
    });
}

int mgk_find_unstable(mgk_ctx* ctx, int* idx_out, double* q_out, int cap, int* n_out) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !n_out) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        const int n = c.hkctl->n_unstable;   // from the last mgk_update_rates
        *n_out = n;
        if (n > 0) {
            kmc_enqueue_unstable(c.kml(), c.stream);
            const int m = std::min(n, cap);
            if (m > 0 && idx_out) MGK_HIP_CHECK(hipMemcpyAsync(idx_out, c.unst_idx, m * sizeof(int), hipMemcpyDeviceToHost, c.stream));
            if (m > 0 && q_out) MGK_HIP_CHECK(hipMemcpyAsync(q_out, c.unst_q, m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        }
        return MGK_OK;
    });
}

int mgk_sync(mgk_ctx* ctx) {
    return guarded([&]() -> int {
        if (!ctx) throw ArgError("null context");
        MGK_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return MGK_OK;
    });
}

int mgk_timer_start(mgk_ctx* ctx) {
    return guarded([&]() -> int {
        if (!ctx) throw ArgError("null context");
        MGK_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
        return MGK_OK;
    });
}

int mgk_timer_stop(mgk_ctx* ctx, double* ms) {
    return guarded([&]() -> int {
        if (!ctx || !ms) throw ArgError("null argument");
        MGK_HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
        MGK_HIP_CHECK(hipEventSynchronize(ctx->ev1));
        float f = 0.f;
        MGK_HIP_CHECK(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
        *ms = f;
        return MGK_OK;
    });
}

int mgk_tuning(mgk_ctx* ctx, int* choices, double* ms) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !choices) throw ArgError("null argument");
        const Geom& g = ctx->g;
        const auto it = g_tunes.find(std::vector<int>{g.dims, g.nx, g.ny, g.nz, ctx->precision});
        choices[0] = ctx->tune.fwd[0];
        choices[1] = ctx->tune.fwd[1];
        choices[2] = ctx->tune.mid[0];
        choices[3] = ctx->tune.mid[1];
        choices[4] = (ctx->precision == MGK_FP32_REFINE && ls_cycle_supported(g)) ? ctx->use_cycle : -1;
        choices[5] = (it != g_tunes.end()) ? it->second.tuned : 0;
        if (ms)
            for (int k = 0; k < 10; ++k) ms[k] = (it != g_tunes.end()) ? it->second.ms[k] : 0.0;
        return MGK_OK;
    });
}

int mgk_set_tuning(mgk_ctx* ctx, const int* choices) {
    return guarded([&]() -> int {
        if (!ctx || !choices) throw ArgError("null argument");
        for (int k = 0; k < 4; ++k)
            if (choices[k] != 0 && choices[k] != 1) throw ArgError("variant must be 0 or 1");
        ctx->tune.fwd[0] = choices[0];
        ctx->tune.fwd[1] = choices[1];
        ctx->tune.mid[0] = choices[2];
        ctx->tune.mid[1] = choices[3];
        ctx->use_cycle = choices[4] != 0 ? 1 : 0;
        return MGK_OK;
    });
}

int mgk_cycle_timer(mgk_ctx* ctx, int on) {
    return guarded([&]() -> int {
        if (!ctx) throw ArgError("null context");
        ctx->cyc_timing = on != 0;
        ctx->cyc_ms = 0.0;
        ctx->cyc_launches = 0;
        ctx->cyc_steps = 0;
        return MGK_OK;
    });
}

int mgk_cycle_timer_read(mgk_ctx* ctx, double* total_ms, long long* launches, long long* steps) {
    return guarded([&]() -> int {
        if (!ctx || !total_ms || !launches || !steps) throw ArgError("null argument");
        *total_ms = ctx->cyc_ms;
        *launches = ctx->cyc_launches;
        *steps = ctx->cyc_steps;
        return MGK_OK;
    });
}

int mgk_cycle_supported(mgk_ctx* ctx, int* on) {
    return guarded([&]() -> int {
        if (!ctx || !on) throw ArgError("null argument");
        *on = (ctx->precision == MGK_FP32_REFINE && ctx->use_cycle && ls_cycle_supported(ctx->g)) ? 1 : 0;
        return MGK_OK;
    });
}

int mgk_time_ls_pass(mgk_ctx* ctx, int pass, int fp32, int reps, double* avg_ms) {
    return guarded([&]() -> int {
        if (!ctx || !avg_ms || reps < 1) throw ArgError("bad argument");
        if (pass < 0 || pass > 11 || (ctx->g.nm == 1 && (pass == 1 || pass == 3))) throw ArgError("bad pass");
        if ((pass == 5 || pass == 6) && (!fp32 || !ctx->corr)) throw ArgError("passes 5 and 6 need an FP32+refine context");
        if (pass >= 6 && pass <= 8 && (!fp32 || !ls_plane_supported(ctx->g)))
            throw ArgError("plane passes: FP32 on a supported 3D grid");
        if (pass >= 9 && (fp32 || !ls_plane64_path(ctx->g))) throw ArgError("FP64 plane passes: FP64 on a supported 3D grid");
        mgk_ctx& c = *ctx;
        // a control block that lets every pass run (not done, shift off)
        std::memset(c.hctl, 0, sizeof(Ctl));
        c.hctl->iters = 1;
        c.hctl->max_iter = 1 << 30;
        MGK_HIP_CHECK(hipMemcpyAsync(c.ctl, c.hctl, sizeof(Ctl), hipMemcpyHostToDevice, c.stream));
        LsParams q = c.ls_params();
        q.corr = (pass == 5 || pass == 6) ? c.corr : nullptr;
        const int mode = fp32 ? LS_CG : LS_FULL;
        q.r32 = fp32 ? c.cg : nullptr;   // the FP32 passes read and write the CG vectors
        const LsLaunch p = c.lsl(1, &q);
        ls_enqueue_pass(p, pass, mode, c.stream);   // warm-up launch
        MGK_HIP_CHECK(hipEventRecord(c.ev0, c.stream));
        for (int r = 0; r < reps; ++r) ls_enqueue_pass(p, pass, mode, c.stream);
        MGK_HIP_CHECK(hipEventRecord(c.ev1, c.stream));
        MGK_HIP_CHECK(hipEventSynchronize(c.ev1));
        float ms = 0.f;
        MGK_HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        *avg_ms = (double)ms / reps;
        return MGK_OK;
    });
}

// ---------------------------------------------------------------------------
// Input generation on the device (gen_kernels.hip)
// ---------------------------------------------------------------------------
int mgk_kmc_draw(mgk_ctx* ctx, const mgk_draw_params* p, unsigned long long seed, int draw_catalog) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !p) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        const long nm = c.g.N * c.kp.M;
        DrawParams d{p->barrier_type, p->b_mean, p->b_std, p->b_min_cutoff, p->b_max_cutoff, p->b_loc, p->b_scale,
                     p->b_epsilon, p->b_ratio, p->gamma0, (c.g.dims == 3 && p->use_3d_barriers) ? 1 : 0};
        double* cat = nullptr;
        if (draw_catalog) {
            c.ensure_cat((size_t)c.g.C * nm);
            cat = c.K.cat;
        }
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        gen_enqueue_modes(c.K.Q0, cat, nm, c.g.C, d, seed, c.stream);
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        if (draw_catalog) {
            c.K.cat_shared = 0;
            c.shared_cat.clear();
        }
        return MGK_OK;
    });
}

int mgk_correlated_field(int dims, int nx, int ny, int nz, double mean, double std, double corr, int has_min,
                         double clip_min, int has_max, double clip_max, unsigned long long seed, const double* noise,
                         double* out) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!out) throw ArgError("null argument");
        if ((dims != 2 && dims != 3) || nx < 1 || ny < 1 || (dims == 3 && nz < 1)) throw ArgError("bad grid");
        const int n[3] = {nx, ny, dims == 3 ? nz : 1};
        FieldParams f{mean, std, corr, has_min ? 1 : 0, has_max ? 1 : 0, clip_min, clip_max};
        gen_correlated_field(n, dims, f, seed, noise, out);
        return MGK_OK;
    });
}

int mgk_correlated_field_noalloc(int dims, int nx, int ny, int nz, double mean, double std, double corr,
                                 int has_min, double clip_min, int has_max, double clip_max, unsigned long long seed,
                                 const double* noise_dev, double* out_dev, void* stream) {
    return guarded([&]() -> int {
        if (!out_dev) throw ArgError("null output");
        if ((dims != 2 && dims != 3) || nx < 1 || ny < 1 || (dims == 3 && nz < 1)) throw ArgError("bad grid");
        const int n[3] = {nx, ny, dims == 3 ? nz : 1};
        FieldParams f{mean, std, corr, has_min ? 1 : 0, has_max ? 1 : 0, clip_min, clip_max};
        gen_correlated_field_noalloc(n, dims, f, seed, noise_dev, out_dev, static_cast<hipStream_t>(stream));
        return MGK_OK;
    });
}

// ---------------------------------------------------------------------------
// Fast patching (kmc_solver.hpp:276-332, :1256-1331, :867-878)
// ---------------------------------------------------------------------------
namespace {
// A scratch FP64 context on the same grid and material, for the unit solves.
mgk_ctx* scratch_like(mgk_ctx& c) {
    const long N = c.g.N;
    std::vector<double> E(N), nu(N);
    MGK_HIP_CHECK(hipMemcpy(E.data(), c.E, N * sizeof(double), hipMemcpyDeviceToHost));
    MGK_HIP_CHECK(hipMemcpy(nu.data(), c.nu, N * sizeof(double), hipMemcpyDeviceToHost));
    mgk_ctx* t = nullptr;
    if (mgk_create(c.g.dims, c.g.nx, c.g.ny, c.g.nz, c.pixel, c.plane_mode, MGK_FP64, E.data(), nu.data(), &t) != MGK_OK)
        throw HipError(std::string("scratch context: ") + mgk_last_error());
    return t;
}
// (i, j) of symmetric component k
void comp_ij(int dims, int k, int& i, int& j) {
    static const int i2[3] = {0, 1, 0}, j2[3] = {0, 1, 1};
    static const int i3[6] = {0, 1, 2, 0, 0, 1}, j3[6] = {0, 1, 2, 1, 2, 2};
    i = dims == 2 ? i2[k] : i3[k];
    j = dims == 2 ? j2[k] : j3[k];
}
}  // namespace

int mgk_patch_init(mgk_ctx* ctx, int radius) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx) throw ArgError("null context");
        mgk_ctx& c = *ctx;
        if (!c.kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        const Geom& g = c.g;
        const bool zwin = g.dims == 3 && g.nz > 1;
        if (radius < 0 || radius > g.nx || radius > g.ny || (zwin && radius > g.nz))
            throw ArgError("patch radius must lie in [0, every grid length]");
        const int D = g.dims, F = D * D, C = g.C, nb = (D == 2) ? 2 : 5;
        const int R = radius, side = 2 * R + 1, sz = zwin ? side : 1, Z = zwin ? R : 0, W = side * side * sz;
        const long N = g.N;
        const int cx = g.nx / 2, cy = g.ny / 2, cz = (D == 3) ? g.nz / 2 : 0;
        auto vid = [&](int x, int y, int z) { return D == 2 ? (long)x * g.ny + y : ((long)x * g.ny + y) * g.nz + z; };
        std::vector<double> K((size_t)nb * C * W), Km((size_t)nb * C);
        mgk_ctx* t = scratch_like(c);
        try {
            std::vector<double> ep((size_t)N * F), e((size_t)N * F), sg((size_t)N * F);
            for (int b = 0; b < nb; ++b) {
                // unit deviatoric basis b at the centre voxel, column-major
                std::fill(ep.begin(), ep.end(), 0.0);
                double* B = &ep[(size_t)vid(cx, cy, cz) * F];
                auto at = [&](int i, int j) -> double& { return B[i + D * j]; };
                if (b == 0) { at(0, 0) = 1.0; at(1, 1) = -1.0; }
                else if (D == 2) { at(0, 1) = 1.0; at(1, 0) = 1.0; }
                else if (b == 1) { at(1, 1) = 1.0; at(2, 2) = -1.0; }
                else if (b == 2) { at(0, 1) = 1.0; at(1, 0) = 1.0; }
                else if (b == 3) { at(0, 2) = 1.0; at(2, 0) = 1.0; }
                else { at(1, 2) = 1.0; at(2, 1) = 1.0; }
                double bar[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, em[9], sm[9], diff;
                int it;
                if (mgk_ls_solve(t, bar, e.data(), 0, ep.data(), sg.data(), em, sm, 200, 1e-6, &it, &diff) != MGK_OK)
                    throw HipError(std::string("patch kernel solve: ") + mgk_last_error());
                // crop the window around the centre (periodic) and the mean it misses
                double sum_crop[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, sum_sig[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int dx = -R; dx <= R; ++dx)
                    for (int dy = -R; dy <= R; ++dy)
                        for (int dz = -Z; dz <= Z; ++dz) {
                            const int sx = (cx + dx + g.nx) % g.nx, sy = (cy + dy + g.ny) % g.ny;
                            const int szz = (D == 3) ? (cz + dz + g.nz) % g.nz : 0;
                            const double* val = &sg[(size_t)vid(sx, sy, szz) * F];
                            const int w = ((dx + R) * side + (dy + R)) * sz + (dz + Z);
                            for (int k = 0; k < C; ++k) {
                                int i, j;
                                comp_ij(D, k, i, j);
                                K[((size_t)b * C + k) * W + w] = val[i + D * j];
                            }
                            for (int f = 0; f < F; ++f) sum_crop[f] += val[f];
                        }
                for (long v = 0; v < N; ++v)
                    for (int f = 0; f < F; ++f) sum_sig[f] += sg[(size_t)v * F + f];
                for (int k = 0; k < C; ++k) {
                    int i, j;
                    comp_ij(D, k, i, j);
                    const int f = i + D * j;
                    Km[(size_t)b * C + k] = sum_sig[f] / N - sum_crop[f] / N;
                }
            }
        } catch (...) {
            mgk_destroy(t);
            throw;
        }
        mgk_destroy(t);
        if (!c.pk) {
            c.pk = c.dal<double>(K.size());
            c.pkm = c.dal<double>(Km.size());
            c.psu = c.dal<double>(6);
        } else if (c.pp.nb * C * c.pp.W != (int)K.size()) {
            MGK_HIP_CHECK(hipFree(c.pk));
            c.dev_bytes -= std::max<size_t>((size_t)c.pp.nb * C * c.pp.W, 1) * sizeof(double);
            c.pk = c.dal<double>(K.size());
        }
        MGK_HIP_CHECK(hipMemcpy(c.pk, K.data(), K.size() * sizeof(double), hipMemcpyHostToDevice));
        MGK_HIP_CHECK(hipMemcpy(c.pkm, Km.data(), Km.size() * sizeof(double), hipMemcpyHostToDevice));
        c.pp = PatchParams{c.pk, c.pkm, c.psu, R, side, sz, Z, nb, W};
        c.patch_ready = true;
        return MGK_OK;
    });
}

int mgk_patch_unit(mgk_ctx* ctx, const double* strain_unit) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !strain_unit) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (!c.patch_ready) throw ArgError("fast patching not initialised (mgk_patch_init)");
        const int D = c.g.dims, C = c.g.C;
        std::vector<double> e((size_t)c.g.N * D * D);
        double em[9], sm[9], diff;
        int it;
        mgk_ctx* t = scratch_like(c);
        const int rc = mgk_ls_solve(t, strain_unit, e.data(), 0, nullptr, nullptr, em, sm, 200, 1e-6, &it, &diff);
        mgk_destroy(t);
        if (rc != MGK_OK) throw HipError(std::string("sigma_macro_unit solve: ") + mgk_last_error());
        double su[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < C; ++k) {
            int i, j;
            comp_ij(D, k, i, j);
            su[k] = (i == j) ? sm[i + D * j] : 0.5 * (sm[i + D * j] + sm[j + D * i]);
        }
        MGK_HIP_CHECK(hipMemcpy(c.psu, su, sizeof(su), hipMemcpyHostToDevice));
        c.unit_ready = true;
        return MGK_OK;
    });
}

int mgk_patch_apply(mgk_ctx* ctx, int voxel, const double* strain, double pred) {
    return guarded([&]() -> int {
        if (!ctx) throw ArgError("null context");
        mgk_ctx& c = *ctx;
        if (!c.patch_ready) throw ArgError("fast patching not initialised (mgk_patch_init)");
        if (pred != 0.0 && !c.unit_ready) throw ArgError("sigma_macro_unit not set (mgk_patch_unit)");
        if (voxel >= c.g.N) throw ArgError("voxel out of range");
        if (voxel >= 0 && !strain) throw ArgError("null strain");
        const int D = c.g.dims;
        double cs[6] = {0, 0, 0, 0, 0, 0};
        if (voxel >= 0)
            for (int k = 0; k < c.g.C; ++k) {
                int i, j;
                comp_ij(D, k, i, j);
                cs[k] = strain[i + D * j];
            }
        kmc_enqueue_patch(c.kml(), c.pp, voxel, cs, pred, 0, 0.0, c.stream);
        return MGK_OK;
    });
}

// ---------------------------------------------------------------------------
// Trajectory state snapshot: header, then the fields in device order.
// ---------------------------------------------------------------------------
namespace {
struct StateHeader {
    char magic[8];
    int version;
    int dims, nx, ny, nz, C, M;
    int cat_shared, epsp_set, mean_valid;
    long long cat_n;                    // doubles of catalog storage
    unsigned long long material_hash;
    double W[3];
    double mean_eps[6];
    int predicted, inner_per_outer, inner_hist[4];
    unsigned inner_pos;
    int last_iters, last_inner;
    int inner_per_outer2, inner_hist2[4];
    unsigned inner_pos2;
};
const char kStateMagic[8] = {'M', 'G', 'K', 'S', 'T', 'A', 'T', 0};

struct StateArray {
    void* dev;
    long long n;   // doubles
};

std::vector<StateArray> state_arrays(mgk_ctx& c, int M, long long cat_n) {
    const long long N = c.g.N, C = c.g.C;
    std::vector<StateArray> a = {{c.eps, C * N}, {c.sig, C * N}, {c.epsp, C * N}};
    if (M > 0) {
        a.push_back({c.K.Q, N * M});
        a.push_back({c.K.Q0, N * M});
        a.push_back({c.K.cat, cat_n});
        a.push_back({c.K.soft, 2 * N});
        a.push_back({c.K.last_t, N});
        a.push_back({c.K.prev_dir, C * N});
    }
    return a;
}

long long cat_doubles(const mgk_ctx& c) {
    if (!c.kmc) return 0;
    return (long long)c.g.C * (c.K.cat_shared ? c.kp.M : c.g.N * c.kp.M);
}
}  // namespace

long long mgk_state_size(const mgk_ctx* ctx) {
    if (!ctx) return -1;
    mgk_ctx& c = const_cast<mgk_ctx&>(*ctx);
    long long b = sizeof(StateHeader);
    for (const StateArray& a : state_arrays(c, c.kmc ? c.kp.M : 0, cat_doubles(c))) b += a.n * (long long)sizeof(double);
    return b;
}

int mgk_state_save(mgk_ctx* ctx, void* buf, long long bytes) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !buf) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        if (bytes < mgk_state_size(ctx)) throw ArgError("state buffer too small (mgk_state_size)");
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        StateHeader h{};
        std::memcpy(h.magic, kStateMagic, sizeof(h.magic));
        h.version = 2;
        h.dims = c.g.dims; h.nx = c.g.nx; h.ny = c.g.ny; h.nz = c.g.nz; h.C = c.g.C;
        h.M = c.kmc ? c.kp.M : 0;
        h.cat_shared = c.kmc ? c.K.cat_shared : 0;
        h.epsp_set = c.epsp_set ? 1 : 0;
        h.mean_valid = c.mean_valid ? 1 : 0;
        h.cat_n = cat_doubles(c);
        h.material_hash = c.material_hash;
        for (int k = 0; k < 3; ++k) h.W[k] = c.W[k];
        for (int k = 0; k < 6; ++k) h.mean_eps[k] = c.mean_eps[k];
        h.predicted = c.predicted; h.inner_per_outer = c.inner_per_outer;
        for (int k = 0; k < 4; ++k) h.inner_hist[k] = c.inner_hist[k];
        h.inner_pos = c.inner_pos; h.last_iters = c.last_iters; h.last_inner = c.last_inner;
        h.inner_per_outer2 = c.inner_per_outer2; h.inner_pos2 = c.inner_pos2;
        for (int k = 0; k < 4; ++k) h.inner_hist2[k] = c.inner_hist2[k];
        char* out = static_cast<char*>(buf);
        std::memcpy(out, &h, sizeof(h));
        out += sizeof(h);
        for (const StateArray& a : state_arrays(c, h.M, h.cat_n)) {
            const size_t nb = (size_t)a.n * sizeof(double);
            MGK_HIP_CHECK(hipMemcpy(out, a.dev, nb, hipMemcpyDeviceToHost));
            out += nb;
        }
        return MGK_OK;
    });
}

int mgk_state_load(mgk_ctx* ctx, const void* buf, long long bytes) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctx || !buf) throw ArgError("null argument");
        mgk_ctx& c = *ctx;
        StateHeader h;
        if (bytes < (long long)sizeof(h)) throw ArgError("state buffer too small");
        std::memcpy(&h, buf, sizeof(h));
        if (std::memcmp(h.magic, kStateMagic, sizeof(h.magic)) != 0 || h.version != 2)
            throw ArgError("not a trajectory state (bad magic or version)");
        if (h.dims != c.g.dims || h.nx != c.g.nx || h.ny != c.g.ny || h.nz != c.g.nz || h.C != c.g.C)
            throw ArgError("state is for another grid");
        if (h.material_hash != c.material_hash) throw ArgError("state is for another material (E, nu)");
        if (h.M > 0 && (!c.kmc || c.kp.M != h.M)) throw ArgError("state holds KMC fields for another mode count");
        if (h.M == 0 && c.kmc) throw ArgError("state holds no KMC fields");
        const long long want = (long long)c.g.C * (h.cat_shared ? h.M : c.g.N * h.M);
        if (h.M > 0 && h.cat_n != want) throw ArgError("state catalog size is inconsistent");
        long long need = sizeof(h);
        if (h.M > 0) c.ensure_cat((size_t)h.cat_n);
        const std::vector<StateArray> arrays = state_arrays(c, h.M, h.cat_n);
        for (const StateArray& a : arrays) need += a.n * (long long)sizeof(double);
        if (bytes < need) throw ArgError("state buffer truncated");
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        const char* in = static_cast<const char*>(buf) + sizeof(h);
        for (const StateArray& a : arrays) {
            const size_t nb = (size_t)a.n * sizeof(double);
            MGK_HIP_CHECK(hipMemcpy(a.dev, in, nb, hipMemcpyHostToDevice));
            in += nb;
        }
        if (h.M > 0) {
            c.K.cat_shared = h.cat_shared;
            if (h.cat_shared) {
                c.shared_cat.resize((size_t)h.cat_n);
                MGK_HIP_CHECK(hipMemcpy(c.shared_cat.data(), c.K.cat, (size_t)h.cat_n * sizeof(double), hipMemcpyDeviceToHost));
            } else {
                c.shared_cat.clear();
            }
        }
        c.epsp_set = h.epsp_set != 0;
        c.mean_valid = h.mean_valid != 0;
        for (int k = 0; k < 3; ++k) c.W[k] = h.W[k];
        for (int k = 0; k < 6; ++k) c.mean_eps[k] = h.mean_eps[k];
        c.predicted = h.predicted; c.inner_per_outer = h.inner_per_outer;
        for (int k = 0; k < 4; ++k) c.inner_hist[k] = h.inner_hist[k];
        c.inner_pos = h.inner_pos; c.last_iters = h.last_iters; c.last_inner = h.last_inner;
        c.inner_per_outer2 = h.inner_per_outer2; c.inner_pos2 = h.inner_pos2;
        for (int k = 0; k < 4; ++k) c.inner_hist2[k] = h.inner_hist2[k];
        c.kctl_pending = false;
        MGK_HIP_CHECK(hipDeviceSynchronize());
        return MGK_OK;
    });
}

long long mgk_device_bytes(const mgk_ctx* ctx) { return ctx ? (long long)ctx->dev_bytes : -1; }

long long mgk_estimate_bytes(int dims, int nx, int ny, int nz, int M, int per_voxel_catalog) {
    if ((dims != 2 && dims != 3) || nx < 1 || ny < 1 || (dims == 3 && nz < 1) || M < 0) return -1;
    const Geom g = make_geom(dims, nx, ny, nz);
    const long long N = g.N, C = g.C, NS = g.NS;
    auto r = [](long long n, long long sz) { return std::max<long long>(n, 1) * sz; };
    long long b = 0;
    b += 4 * r(N, 8) + 2 * r(N, 4);                      // E, nu, lam, mu; dlam, dmu
    b += 3 * r(C * N, 8);                                 // eps, sig, eps_p
    b += r(2 * C * NS, 8) + r(4 * C * N, 4) + r(C * N, 4);   // spectrum, CG vectors, correction
    b += r((long long)ls_partial_slots(g) + 64, 8);
    b += (long long)sizeof(CtlPair) + 2 * (long long)sizeof(LsParams) + (long long)sizeof(KmcParams) +
         (long long)sizeof(EventArgs);
    const int lens[3] = {g.nf, g.nm, g.nl};
    for (int a = 0; a < 3; ++a) b += r(2LL * lens[a], 8) + r(2LL * lens[a], 4);   // twiddles
    if (M > 0) {
        const long long nm = N * M, nb = kmc_nblocks(nm);
        b += 2 * r(nm, 8);                                // Q, Q0
        b += per_voxel_catalog ? r(C * nm, 8) : r(C * M, 8);
        b += r(2 * N, 8) + r(N, 8) + r(C * N, 8);         // softening, last event time, previous direction
        b += 2 * r(nb, 8) + 3 * r(nb, 4);                 // block sums and counts
        b += r((long long)M * (C + 1), 8);
        b += r((long long)kFlipChunk * M * (C + 1), 8) + r(kFlipChunk, 4);   // cascade flip lists
        b += r(nm, 4) + r(nm, 8);                         // unstable list
    }
    return b;
}

int mgk_last_solve_stats(mgk_ctx* ctx, int* fp64_iters, int* fp32_iters) {
    return guarded([&]() -> int {
        if (!ctx) throw ArgError("null context");
        if (fp64_iters) *fp64_iters = ctx->last_iters;
        if (fp32_iters) *fp32_iters = ctx->last_inner;
        return MGK_OK;
    });
}

int mgk_last_solve_history(mgk_ctx* ctx, double* diffs, int cap, int* n_out) {
    return guarded([&]() -> int {
        if (!ctx || !n_out || (cap > 0 && !diffs)) throw ArgError("null argument");
        const Ctl& h = *ctx->hctl;
        const int n = std::min(32, h.iters + h.inner);
        *n_out = n;
        for (int k = 0; k < std::min(n, cap); ++k) diffs[k] = h.diff_hist[k];
        return MGK_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// A batch of trajectories of one shape that advance together: every launch
// of a batch event covers all of them (block row y of each grid works on
// trajectory y, through its own parameter block).  The contexts keep their
// own fields and host state, and stay usable on their own between batch
// calls (cascades, field reads).
// ---------------------------------------------------------------------------
struct mgk_batch {
    std::vector<mgk_ctx*> c;
    int B = 0;
    hipStream_t stream = nullptr;
    LsParams *d_lsp = nullptr, *h_lsp = nullptr;
    KmcParams *d_kp = nullptr, *h_kp = nullptr;
    EventArgs *d_ea = nullptr, *h_ea = nullptr;
    double *d_times = nullptr, *h_times = nullptr;
    Ctl *d_ctl = nullptr, *h_ctl = nullptr;
    KmcCtl *d_kctl = nullptr, *h_kctl = nullptr;
    double *d_stage = nullptr, *h_stage = nullptr;
    size_t stage_per = 0;   // M * (C + 1) doubles per trajectory

    // The first round of an event call (uploads from the pinned buffers, event
    // kernels, the predicted iterations, final pass, read-back) as HIP
    // graphs, one per launch structure.  Every per-call value is read from
    // device memory the graph's own copies fill, so a graph replays as is.
    // MGK_BATCH_GRAPHS=0 starts batches with eager launches (the same kernels
    // in the same order), for tools that cannot follow stream capture.
    bool use_graphs = !(std::getenv("MGK_BATCH_GRAPHS") && std::getenv("MGK_BATCH_GRAPHS")[0] == '0');
    std::map<std::vector<int>, hipGraphExec_t> graphs;

    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    ~mgk_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        auto f = [](void* p) { if (p) (void)hipFree(p); };
        auto fh = [](void* p) { if (p) (void)hipHostFree(p); };
        f(d_lsp); f(d_kp); f(d_ea); f(d_times); f(d_ctl); f(d_kctl); f(d_stage);
        fh(h_lsp); fh(h_kp); fh(h_ea); fh(h_times); fh(h_ctl); fh(h_kctl); fh(h_stage);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// One event attempt for every active trajectory of the batch, the same
// sequence of launches as mgk_kmc_event with grids B rows tall.
void batch_event(mgk_batch& b, const mgk_event_in* in, const int* active, double* eps_macro, double* sig_macro,
                 mgk_event_out* out) {
    const int B = b.B;
    mgk_ctx& c0 = *b.c[0];
    const Geom g = c0.g;
    const int Dd = c0.D(), F = Dd * Dd, C = g.C;
    const bool refine = (c0.precision == MGK_FP32_REFINE);
    int max_iter = 0, J = 1, J2 = 1, predicted = 1, sp = 1 << 30, sp2 = 1 << 30;
    double tol = 0.0;
    int first_mode = 0, n_modes = -1, has_cat = 0, has_q0 = 0;
    bool all_keep = true;
    int n_active = 0;
    for (int y = 0; y < B; ++y) {
        mgk_ctx& c = *b.c[y];
        const mgk_event_in& e = in[y];
        const bool on = !active || active[y];
        // the contexts' own streams hold no pending work behind a batch call
        MGK_HIP_CHECK(hipStreamSynchronize(c.stream));
        c.kctl_pending = false;
        if (on) {
            if (e.drv_i < 0 || e.drv_i >= Dd || e.drv_j < 0 || e.drv_j >= Dd) throw ArgError("driven component out of range");
            if (e.max_iter < 1) throw ArgError("max_iter must be >= 1");
            if (e.patch_mode) throw ArgError("fast patching runs on single trajectories (mgk_kmc_event)");
            int hc = 0, hq = 0;
            const int n = stage_redraw(c, e.new_catalog, e.new_Q0, e.first_mode, e.n_modes, &hc, &hq,
                                       b.h_stage + (size_t)y * b.stage_per);
            if (n_active == 0) {
                max_iter = e.max_iter; tol = e.tol; first_mode = e.first_mode;
                n_modes = n; has_cat = hc; has_q0 = hq;
            } else if (e.max_iter != max_iter || e.tol != tol || e.first_mode != first_mode || n != n_modes ||
                       hc != has_cat || hq != has_q0) {
                throw ArgError("the trajectories of a batch share max_iter, tol and the redrawn modes");
            }
            ++n_active;
        }
        // LS control block, parameter blocks and event arguments; an inactive
        // trajectory's host state stays as it is (its row's control block is
        // done before the first iteration)
        Ctl& h = b.h_ctl[y];
        if (on) {
            c.init_ctl(e.max_iter, e.tol);
            h = *c.hctl;
        } else {
            std::memset(&h, 0, sizeof(h));
            h.max_iter = 1;
            h.done = 1;
            h.skip_final = 1;
        }
        h.cycle_len = c.inner_per_outer;   // this trajectory's own cycles, as when it runs alone
        h.cycle_len2 = c.inner_per_outer2;
        const int keep = c.mean_valid ? 1 : 0;
        if (on) all_keep = all_keep && keep;
        h.keep_shift = keep;
        LsParams lp = c.ls_params();
        lp.ctl = b.d_ctl + y;
        lp.gbar = nullptr;   // batches run their CG steps as plane passes
        b.h_lsp[y] = lp;
        KmcParams kp = c.K;
        kp.kctl = b.d_kctl + y;
        kp.staging = b.d_stage + (size_t)y * b.stage_per;
        kp.skip = on ? 0 : 1;
        b.h_kp[y] = kp;
        EventArgs& ea = b.h_ea[y];
        std::memset(&ea, 0, sizeof(ea));
        if (!on) {
            ea.n_vox = (double)g.N;
            ea.active = 0;
            b.h_times[y] = 0.0;
            continue;
        }
        ea.time = e.time;
        ea.neg_log_u1 = e.neg_log_u1;
        ea.remaining = e.remaining;
        ea.strain_rate = e.strain_rate;
        for (int k = 0; k < F; ++k) ea.a[k] = eps_macro[(size_t)y * F + k];
        ea.n_vox = (double)g.N;
        ea.u2 = e.u2;
        ea.drv = e.drv_i + Dd * e.drv_j;
        ea.cascade_check = e.cascade_check ? 1 : 0;
        ea.keep_shift = keep;
        ea.active = on ? 1 : 0;
        b.h_times[y] = on ? e.time : 0.0;
        J = std::max(J, c.inner_per_outer);
        J2 = std::max(J2, c.inner_per_outer2);
        sp = std::min(sp, c.spec_from());
        sp2 = std::min(sp2, c.spec_from2());
        predicted = std::max(predicted, c.predicted);
    }
    if (n_active == 0) throw ArgError("no active trajectory in the batch call");
    hipStream_t s = b.stream;
    const KmcLaunch KL{g, c0.kp.M, b.d_kp, B, b.h_kp[0], true};
    LsLaunch L{g, b.d_lsp, B, b.h_lsp[0], true, ls_plane_path(g)};
    L.plane64 = ls_plane64_path(g);
    L.tune = c0.tune;   // every member has the batch's shape, hence its tuning
    const int nb0 = std::min(max_iter, std::max(refine ? 1 : 2, predicted));
    auto iterations = [&](int from, int count, hipStream_t st) {
        for (int q = 0; q < count; ++q) {
            const int k = from + q;
            const int flags = (k == 0 ? LS_FLAG_SHIFT : 0) | ((refine && k > 0) ? LS_FLAG_CORR : 0);
            ls_enqueue_iteration(L, LS_FULL, flags, st);
            if (refine) {
                const int Jk = (k == 0) ? J : J2, spk = (k == 0) ? sp : sp2;
                for (int j = 0; j < Jk; ++j) ls_enqueue_iteration(L, LS_CG, mgk_ctx::cg_flags(j, spk, Jk), st);
            }
        }
        ls_enqueue_final(L, st);
        MGK_HIP_CHECK(hipMemcpyAsync(b.h_ctl, b.d_ctl, sizeof(Ctl) * B, hipMemcpyDeviceToHost, st));
        MGK_HIP_CHECK(hipMemcpyAsync(b.h_kctl, b.d_kctl, sizeof(KmcCtl) * B, hipMemcpyDeviceToHost, st));
    };
    auto first_round = [&](hipStream_t st) {
        MGK_HIP_CHECK(hipMemcpyAsync(b.d_lsp, b.h_lsp, sizeof(LsParams) * B, hipMemcpyHostToDevice, st));
        MGK_HIP_CHECK(hipMemcpyAsync(b.d_kp, b.h_kp, sizeof(KmcParams) * B, hipMemcpyHostToDevice, st));
        MGK_HIP_CHECK(hipMemcpyAsync(b.d_ea, b.h_ea, sizeof(EventArgs) * B, hipMemcpyHostToDevice, st));
        MGK_HIP_CHECK(hipMemcpyAsync(b.d_times, b.h_times, sizeof(double) * B, hipMemcpyHostToDevice, st));
        MGK_HIP_CHECK(hipMemcpyAsync(b.d_ctl, b.h_ctl, sizeof(Ctl) * B, hipMemcpyHostToDevice, st));
        if (n_modes > 0)
            MGK_HIP_CHECK(hipMemcpyAsync(b.d_stage, b.h_stage, sizeof(double) * b.stage_per * B, hipMemcpyHostToDevice, st));
        kmc_enqueue_rates(KL, 0.0, b.d_times, st, false);
        kmc_enqueue_event_chain(KL, nullptr, b.d_lsp, b.h_ea[0], b.d_ea, refine, first_mode, std::max(n_modes, 0),
                                has_cat, has_q0, nullptr, st);
        if (!refine) {
            kmc_enqueue_select_fallback(KL, 1, st);
            kmc_enqueue_flip(KL, -1, 0.0, first_mode, std::max(n_modes, 0), has_cat, has_q0, st, 1);
        }
        if (!all_keep) ls_enqueue_start(L, LS_START_MEAN, st);   // trajectories with a known mean keep their shift
        iterations(0, nb0, st);
    };
    if (b.use_graphs) {
        const std::vector<int> key{J,       nb0,         all_keep ? 1 : 0, n_modes, first_mode, has_cat, has_q0,
                                   refine ? 1 : 0, std::min(sp, J), J2, std::min(sp2, J2)};
        auto it = b.graphs.find(key);
        if (it == b.graphs.end()) {
            std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
            if (b.graphs.size() >= 64) b.drop_graphs();
            hipGraph_t graph = nullptr;
            MGK_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                first_round(s);
            } catch (...) {
                hipGraph_t partial = nullptr;
                (void)hipStreamEndCapture(s, &partial);
                if (partial) (void)hipGraphDestroy(partial);
                throw;
            }
            MGK_HIP_CHECK(hipStreamEndCapture(s, &graph));
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            MGK_HIP_CHECK(e);
            it = b.graphs.emplace(key, exec).first;
        }
        MGK_HIP_CHECK(hipGraphLaunch(it->second, s));
    } else {
        first_round(s);
    }
    MGK_HIP_CHECK(hipStreamSynchronize(s));
    // further iterations (eager) until every trajectory's control block says done
    int enqueued = nb0;
    int nb = nb0;
    while (true) {
        bool all_done = true;
        for (int y = 0; y < B; ++y) all_done = all_done && b.h_ctl[y].done;
        if (all_done || enqueued >= max_iter) break;
        nb = std::min(max_iter - enqueued, std::max(refine ? 1 : 2, nb / 2));
        iterations(enqueued, nb, s);
        enqueued += nb;
        MGK_HIP_CHECK(hipStreamSynchronize(s));
    }
    for (int y = 0; y < B; ++y) {
        mgk_ctx& c = *b.c[y];
        const KmcCtl& k = b.h_kctl[y];
        mgk_event_out& o = out[y];
        std::memset(&o, 0, sizeof(o));
        if (active && !active[y]) {
            o.outcome = MGK_EVENT_SKIPPED;
            o.flat_idx = -1;
            continue;
        }
        *c.hctl = b.h_ctl[y];
        *c.hkctl = k;
        o.outcome = k.outcome;
        o.total_rate = k.total_rate;
        o.n_unstable = k.n_unstable;
        o.n_included = k.n_included;
        o.t_wait = k.t_wait;
        o.flat_idx = k.go ? k.sel : -1;
        o.q_sel = k.go ? k.q_sel : 0.0;
        o.exact_selection = (k.go && k.ambiguous && !refine) ? 1 : 0;
        if (k.outcome != MGK_EVENT_DONE) continue;
        c.absorb_solve(tol);
        o.iters = c.hctl->iters;
        o.diff = c.hctl->diff;
        c.event_macros(in[y], k.t_wait, eps_macro + (size_t)y * F, sig_macro + (size_t)y * F);
    }
}

}  // namespace

extern "C" {

int mgk_batch_create(mgk_ctx* const* ctxs, int B, mgk_batch** out) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!ctxs || !out || B < 1) throw ArgError("bad argument");
        *out = nullptr;
        const mgk_ctx* a = ctxs[0];
        if (!a) throw ArgError("null context");
        for (int y = 0; y < B; ++y) {
            const mgk_ctx* c = ctxs[y];
            if (!c) throw ArgError("null context");
            if (!c->kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
            if (c->g.dims != a->g.dims || c->g.nx != a->g.nx || c->g.ny != a->g.ny || c->g.nz != a->g.nz ||
                c->precision != a->precision || c->kp.M != a->kp.M)
                throw ArgError("the trajectories of a batch share shape, precision and mode count");
            for (int z = 0; z < y; ++z)
                if (ctxs[z] == c) throw ArgError("a context appears twice in a batch");
        }
        mgk_batch* b = new mgk_batch();
        try {
            b->c.assign(ctxs, ctxs + B);
            b->B = B;
            b->stage_per = (size_t)a->kp.M * (a->g.C + 1);
            MGK_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            b->d_lsp = dalloc<LsParams>(B); b->h_lsp = halloc<LsParams>(B);
            b->d_kp = dalloc<KmcParams>(B); b->h_kp = halloc<KmcParams>(B);
            b->d_ea = dalloc<EventArgs>(B); b->h_ea = halloc<EventArgs>(B);
            b->d_times = dalloc<double>(B); b->h_times = halloc<double>(B);
            b->d_ctl = dalloc<Ctl>(B); b->h_ctl = halloc<Ctl>(B);
            b->d_kctl = dalloc<KmcCtl>(B); b->h_kctl = halloc<KmcCtl>(B);
            b->d_stage = dalloc<double>(b->stage_per * B); b->h_stage = halloc<double>(b->stage_per * B);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
        return MGK_OK;
    });
}

int mgk_batch_destroy(mgk_batch* b) {
    if (!b) return MGK_OK;
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        delete b;
        return MGK_OK;
    });
}

int mgk_batch_size(const mgk_batch* b) { return b ? b->B : 0; }

int mgk_batch_set_member(mgk_batch* b, int slot, mgk_ctx* ctx) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!b || !ctx) throw ArgError("null argument");
        if (slot < 0 || slot >= b->B) throw ArgError("slot out of range");
        if (!ctx->kmc) throw ArgError("KMC state not initialised (mgk_kmc_init)");
        const mgk_ctx* a = b->c[slot == 0 && b->B > 1 ? 1 : 0];
        if (ctx->g.dims != a->g.dims || ctx->g.nx != a->g.nx || ctx->g.ny != a->g.ny || ctx->g.nz != a->g.nz ||
            ctx->precision != a->precision || ctx->kp.M != a->kp.M)
            throw ArgError("the trajectories of a batch share shape, precision and mode count");
        for (int y = 0; y < b->B; ++y)
            if (y != slot && b->c[y] == ctx) throw ArgError("a context appears twice in a batch");
        // the rows' parameter blocks are rebuilt from the members on every
        // event call (and copied by the captured graphs): no graph changes
        MGK_HIP_CHECK(hipStreamSynchronize(b->stream));
        b->c[slot] = ctx;
        return MGK_OK;
    });
}

int mgk_batch_set_graphs(mgk_batch* b, int on) {
    return guarded([&]() -> int {
        std::lock_guard<std::recursive_mutex> device_lock(g_device_mu);
        if (!b) throw ArgError("null batch");
        MGK_HIP_CHECK(hipStreamSynchronize(b->stream));
        if (!on) b->drop_graphs();
        b->use_graphs = on != 0;
        return MGK_OK;
    });
}

int mgk_batch_kmc_event(mgk_batch* b, const mgk_event_in* in, const int* active, double* eps_macro, double* sig_macro,
                        mgk_event_out* out) {
    return guarded([&]() -> int {
        if (!b || !in || !eps_macro || !sig_macro || !out) throw ArgError("null argument");
        batch_event(*b, in, active, eps_macro, sig_macro, out);
        return MGK_OK;
    });
}

}  // extern "C"
