// C ABI of the training-side FP64 builders (whitening, drift interpolation, wPCA).
//
// These run once per recording (or once per batch for the drift matrix), so every entry point is
// synchronous: it takes its scratch from a per-device arena kept between calls, enqueues its
// kernels on the given stream and waits for them.  Inputs and outputs named d_* are device memory
// of the current device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/lss.h"
#include "lss_kernels.h"

namespace lss {
lss_status set_error(lss_status st, const char* msg);
}

namespace {

lss_status err(lss_status st, const std::string& m) { return lss::set_error(st, m.c_str()); }

#define TRY_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return err(LSS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device scratch of one call, carved from a per-device arena that grows to the largest request
// and is kept for later calls (lss_train_release frees it).  A call holds the arena's lock from
// commit() to its end, so concurrent calls on one device run one after the other.
struct Arena {
    void* p = nullptr;
    size_t cap = 0;
};
std::mutex g_arena_mu;
std::map<int, Arena> g_arena;

struct Scratch {
    hipStream_t s;
    std::vector<std::pair<void**, size_t>> req;
    std::unique_lock<std::mutex> lock;
    explicit Scratch(hipStream_t st) : s(st) {}
    template <typename T>
    void add(T** p, size_t count) {
        req.emplace_back((void**)p, (sizeof(T) * std::max<size_t>(count, 1) + 255) / 256 * 256);
    }
    hipError_t commit() {
        size_t total = 0;
        for (auto& r : req) total += r.second;
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        lock = std::unique_lock<std::mutex>(g_arena_mu);
        Arena& a = g_arena[dev];
        if (a.cap < total) {
            if (a.p) {
                if ((e = hipDeviceSynchronize()) != hipSuccess) return e;
                (void)hipFree(a.p);
                a.p = nullptr;
                a.cap = 0;
            }
            if ((e = hipMalloc(&a.p, total)) != hipSuccess) return e;
            a.cap = total;
        }
        size_t off = 0;
        for (auto& r : req) {
            *r.first = (char*)a.p + off;
            off += r.second;
        }
        return hipSuccess;
    }
    ~Scratch() {
        if (lock.owns_lock()) (void)hipStreamSynchronize(s);
    }
};

// out = A . B for [C][C] row-major FP64 matrices (out may alias neither input)
lss_status gemm_square(const double* A, int64_t lda, const double* B, int64_t ldb, double* out, int64_t ldo, int n,
                       double* P, int S, int64_t kchunk, hipStream_t s) {
    TRY_HIP(lss::launch_gemm64_kn(A, lda, B, ldb, P, n, n, n, S, kchunk, s));
    TRY_HIP(lss::launch_reduce64(P, S, n, n, 0, 1.0, 0.0, 0.0, out, ldo, s));
    return LSS_OK;
}

// Linear interpolation with linear extrapolation (scipy.interpolate.interp1d(kind='linear',
// fill_value='extrapolate')): x sorted stably, interval from searchsorted(side='left') clipped to
// [1, n-1], y = slope * (x_new - x_lo) + y_lo.
void interp_linear(const double* x, const double* y, int n, const double* xn, int m, double* out) {
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return x[a] < x[b]; });
    std::vector<double> xs(n), ys(n);
    for (int i = 0; i < n; ++i) { xs[i] = x[ord[i]]; ys[i] = y[ord[i]]; }
    for (int j = 0; j < m; ++j) {
        int idx = (int)(std::lower_bound(xs.begin(), xs.end(), xn[j]) - xs.begin());
        idx = std::min(std::max(idx, 1), n - 1);
        const int lo = idx - 1, hi = idx;
        const double slope = (ys[hi] - ys[lo]) / (xs[hi] - xs[lo]);
        out[j] = slope * (xn[j] - xs[lo]) + ys[lo];
    }
}

// ---- isolated-peak snippets: argument checks and the kernels, over scratch the caller provides

lss_status snippet_args(const char* fn, const float* d_x_wc, int64_t ldx, int32_t C, int64_t W, int32_t nt,
                        int32_t twav_min, const int32_t* loc_range, const int32_t* long_range, const float* d_clips,
                        int64_t max_clips, const int64_t* n_found) {
    if (!d_x_wc || !loc_range || !long_range || !n_found || C < 1 || ldx < C || W < 1 || nt < 1 ||
        twav_min < 0 || twav_min > nt || max_clips < 0 || (max_clips > 0 && !d_clips))
        return err(LSS_ERR_ARG, std::string(fn) + ": bad argument");
    if (loc_range[0] < 0 || loc_range[1] < 0 || loc_range[0] > 8 || loc_range[1] > 8 || long_range[0] < 0 ||
        long_range[1] < 0)
        return err(LSS_ERR_ARG, std::string(fn) + ": loc_range entries must be in [0, 8], long_range >= 0");
    return LSS_OK;
}

// Scratch sizes: the peak and isolated-peak bitmaps hold nwords 32-bit words each (two words per
// 64 samples of each channel), the block-sum scan nb ints, the total one long long.  The kernels
// write every word of all four before reading it, so nothing has to start at zero.
struct SnippetScratch {
    int64_t nw, nwords;
    int nb;
    SnippetScratch(int32_t C, int64_t W)
        : nw((W + 63) / 64 * 2), nwords((int64_t)C * nw), nb(lss::snippet_scan_blocks(nwords)) {}
};

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

lss_status snippet_run(const float* d_x_wc, int64_t ldx, int32_t C, int64_t W, int32_t nt, int32_t twav_min,
                       float th_single_ch, const int32_t* loc_range, const int32_t* long_range, float* d_clips,
                       int64_t max_clips, int32_t* d_chan, int32_t* d_time, int64_t* n_found, const SnippetScratch& z,
                       unsigned* pk, unsigned* iso, int* bsum, long long* total, hipStream_t s) {
    TRY_HIP(lss::launch_isolated_peaks(d_x_wc, ldx, C, W, loc_range[0], loc_range[1], th_single_ch, long_range[0],
                                       long_range[1], nt, pk, iso, z.nw, s));
    TRY_HIP(lss::launch_emit_snippets(iso, z.nwords, z.nw, bsum, total, d_x_wc, ldx, nt, twav_min, d_clips, max_clips,
                                      d_chan, d_time, s));
    long long n = 0;
    TRY_HIP(hipMemcpyAsync(&n, total, sizeof(long long), hipMemcpyDeviceToHost, s));
    TRY_HIP(hipStreamSynchronize(s));
    *n_found = n;
    return LSS_OK;
}

// Fixed scratch of lss_extract_snippets_noalloc.  Each GPU holds its own copy of this array from
// the module's load until the process ends; the library never allocates or frees it.  Host code
// uses the device address from hipGetSymbolAddress (kept per device), never the array's name.
static __device__ __align__(256) unsigned char scratch[LSS_NOALLOC_SCRATCH_BYTES];
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu[kNoAllocDevices];      // one call at a time per device
void* g_noalloc_ptr[kNoAllocDevices] = {};     // scratch's address on each device, once looked up

// With the device `dev` current and g_noalloc_mu[dev] held: scratch's address on that device.
lss_status noalloc_scratch(int dev, char** base) {
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return err(LSS_ERR_HIP, "lss_extract_snippets_noalloc: no device address for the scratch array");
        g_noalloc_ptr[dev] = p;
    }
    *base = (char*)g_noalloc_ptr[dev];
    return LSS_OK;
}

}  // namespace

extern "C" {

lss_status lss_covariance_add(const float* d_x, int64_t ldx, int32_t C, int64_t n, double* d_cc, int64_t ldcc,
                              void* hip_stream) {
    if (!d_x || !d_cc || C < 1 || n < 1 || ldx < n || ldcc < C)
        return err(LSS_ERR_ARG, "lss_covariance_add: bad argument");
    hipStream_t s = (hipStream_t)hip_stream;
    int S;
    int64_t kchunk;
    lss::gemm64_split(C, C, n, 1, &S, &kchunk);
    Scratch sc(s);
    double* P;
    sc.add(&P, (size_t)S * C * C);
    TRY_HIP(sc.commit());
    TRY_HIP(lss::launch_gemm64_f32_nk(d_x, ldx, d_x, ldx, P, C, C, n, S, kchunk, 1, s));
    TRY_HIP(lss::launch_reduce64(P, S, C, C, 1, 1.0 / (double)n, 1.0, 0.0, d_cc, ldcc, s));
    TRY_HIP(hipStreamSynchronize(s));
    return LSS_OK;
}

lss_status lss_whitening_from_covariance(const double* d_cc, int64_t ldcc, int32_t n, double eps, double* d_w,
                                         int64_t ldw, void* hip_stream) {
    if (!d_cc || !d_w || n < 1 || n > lss::kEigMax || ldcc < n || ldw < n)
        return err(LSS_ERR_ARG, "lss_whitening_from_covariance: bad argument (n <= 96)");
    hipStream_t s = (hipStream_t)hip_stream;
    TRY_HIP(lss::launch_whiten_full(d_cc, ldcc, n, eps, d_w, ldw, s));
    TRY_HIP(hipStreamSynchronize(s));
    return LSS_OK;
}

lss_status lss_whitening_local(const double* d_cc, int64_t ldcc, int32_t C, const double* xc, const double* yc,
                               int32_t nrange, double eps, double* d_w, int64_t ldw, void* hip_stream) {
    if (!d_cc || !d_w || !xc || !yc || C < 1 || C > 1024 || nrange < 1 || ldcc < C || ldw < C)
        return err(LSS_ERR_ARG, "lss_whitening_local: bad argument");
    const int nr = std::min<int>(nrange, C);
    if (nr > lss::kEigMax) return err(LSS_ERR_ARG, "lss_whitening_local: nrange must be <= 96");
    hipStream_t s = (hipStream_t)hip_stream;
    std::vector<int> idx((size_t)C * nr);
    std::vector<int> ord(C);
    std::vector<double> ds(C);
    for (int j = 0; j < C; ++j) {
        for (int i = 0; i < C; ++i) {
            const double dx = xc[j] - xc[i], dy = yc[j] - yc[i];
            ds[i] = dx * dx + dy * dy;
        }
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return ds[a] < ds[b]; });
        std::copy(ord.begin(), ord.begin() + nr, idx.begin() + (size_t)j * nr);
    }
    Scratch sc(s);
    int* d_idx;
    sc.add(&d_idx, idx.size());
    TRY_HIP(sc.commit());
    TRY_HIP(hipMemcpyAsync(d_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, s));
    TRY_HIP(lss::launch_whiten_local(d_cc, ldcc, d_idx, nr, C, eps, d_w, ldw, s));
    TRY_HIP(hipStreamSynchronize(s));
    return LSS_OK;
}

lss_status lss_drift_inverse(const double* xc, const double* yc, int32_t C, double sig_interp, double lam,
                             double* d_ikxx, int64_t ldk, int32_t* iterations, void* hip_stream) {
    if (!xc || !yc || !d_ikxx || C < 1 || ldk < C || !(sig_interp > 0.0) || !(lam >= 0.0))
        return err(LSS_ERR_ARG, "lss_drift_inverse: bad argument");
    hipStream_t s = (hipStream_t)hip_stream;
    int S;
    int64_t kchunk;
    lss::gemm64_split(C, C, C, 0, &S, &kchunk);
    const size_t CC = (size_t)C * C;
    Scratch sc(s);
    double *dxc, *dyc, *A, *X, *R, *P, *nrm;
    sc.add(&dxc, C);
    sc.add(&dyc, C);
    sc.add(&A, CC);
    sc.add(&X, CC);
    sc.add(&R, CC);
    sc.add(&P, (size_t)S * CC);
    sc.add(&nrm, 1);
    TRY_HIP(sc.commit());
    TRY_HIP(hipMemcpyAsync(dxc, xc, sizeof(double) * C, hipMemcpyHostToDevice, s));
    TRY_HIP(hipMemcpyAsync(dyc, yc, sizeof(double) * C, hipMemcpyHostToDevice, s));
    TRY_HIP(lss::launch_interp_kernel(dxc, dyc, nullptr, C, sig_interp, lam, A, s));
    // X0 = I / ||A||_inf: A is symmetric positive definite with positive entries, so the
    // eigenvalues of I - A X0 lie in [0, 1) and Newton-Schulz X <- X + X (I - A X) converges.
    std::vector<double> hA(CC);
    TRY_HIP(hipMemcpyAsync(hA.data(), A, sizeof(double) * CC, hipMemcpyDeviceToHost, s));
    TRY_HIP(hipStreamSynchronize(s));
    double ninf = 0.0;
    for (int i = 0; i < C; ++i) {
        double r = 0.0;
        for (int j = 0; j < C; ++j) r += std::fabs(hA[(size_t)i * C + j]);
        ninf = std::max(ninf, r);
    }
    TRY_HIP(lss::launch_reduce64(nullptr, 0, C, C, 0, 0.0, 0.0, 1.0 / ninf, X, C, s));
    double prev = INFINITY, r = INFINITY;
    int it = 0;
    for (; it < 200; ++it) {
        TRY_HIP(lss::launch_gemm64_kn(A, C, X, C, P, C, C, C, S, kchunk, s));
        TRY_HIP(lss::launch_reduce64(P, S, C, C, 0, -1.0, 0.0, 1.0, R, C, s));     // R = I - A X
        TRY_HIP(lss::launch_sumsq64(R, C, C, C, nrm, s));
        double r2 = 0.0;
        TRY_HIP(hipMemcpyAsync(&r2, nrm, sizeof(double), hipMemcpyDeviceToHost, s));
        TRY_HIP(hipStreamSynchronize(s));
        r = std::sqrt(r2);
        // stop at FP64 resolution: the residual no longer shrinks quadratically
        if (r < 1e-300 || (r < 1e-6 && r > 0.5 * prev)) break;
        prev = r;
        TRY_HIP(lss::launch_gemm64_kn(X, C, R, C, P, C, C, C, S, kchunk, s));
        TRY_HIP(lss::launch_reduce64(P, S, C, C, 0, 1.0, 1.0, 0.0, X, C, s));     // X += X R
    }
    if (!(r < 1e-6))
        return err(LSS_ERR_ARG, "lss_drift_inverse: Newton-Schulz did not converge (||I - A X|| = " +
                                    std::to_string(r) + ")");
    TRY_HIP(hipMemcpy2DAsync(d_ikxx, sizeof(double) * ldk, X, sizeof(double) * C, sizeof(double) * C, C,
                             hipMemcpyDeviceToDevice, s));
    TRY_HIP(hipStreamSynchronize(s));
    if (iterations) *iterations = it;
    return LSS_OK;
}

lss_status lss_drift_matrix(const double* xc, const double* yc, int32_t C, const double* yblk, const double* dshift,
                            int32_t nblocks, double sig_interp, const double* d_ikxx, int64_t ldk, double* d_m,
                            int64_t ldm, void* hip_stream) {
    if (!xc || !yc || !dshift || !d_ikxx || !d_m || C < 1 || nblocks < 1 || ldk < C || ldm < C ||
        !(sig_interp > 0.0) || (nblocks > 1 && !yblk))
        return err(LSS_ERR_ARG, "lss_drift_matrix: bad argument");
    hipStream_t s = (hipStream_t)hip_stream;
    std::vector<double> shift(C);
    if (nblocks == 1)
        std::fill(shift.begin(), shift.end(), dshift[0]);
    else
        interp_linear(yblk, dshift, nblocks, yc, C, shift.data());
    int S;
    int64_t kchunk;
    lss::gemm64_split(C, C, C, 0, &S, &kchunk);
    const size_t CC = (size_t)C * C;
    Scratch sc(s);
    double *dxc, *dyc, *dsh, *K, *P;
    sc.add(&dxc, C);
    sc.add(&dyc, C);
    sc.add(&dsh, C);
    sc.add(&K, CC);
    sc.add(&P, (size_t)S * CC);
    TRY_HIP(sc.commit());
    TRY_HIP(hipMemcpyAsync(dxc, xc, sizeof(double) * C, hipMemcpyHostToDevice, s));
    TRY_HIP(hipMemcpyAsync(dyc, yc, sizeof(double) * C, hipMemcpyHostToDevice, s));
    TRY_HIP(hipMemcpyAsync(dsh, shift.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
    TRY_HIP(lss::launch_interp_kernel(dxc, dyc, dsh, C, sig_interp, 0.0, K, s));   // Kyx
    if (lss_status st = gemm_square(K, C, d_ikxx, ldk, d_m, ldm, C, P, S, kchunk, s)) return st;
    TRY_HIP(hipStreamSynchronize(s));
    return LSS_OK;
}

lss_status lss_extract_snippets(const float* d_x_wc, int64_t ldx, int32_t C, int64_t W, int32_t nt, int32_t twav_min,
                                float th_single_ch, const int32_t* loc_range, const int32_t* long_range,
                                float* d_clips, int64_t max_clips, int32_t* d_chan, int32_t* d_time, int64_t* n_found,
                                void* hip_stream) {
    if (lss_status st = snippet_args("lss_extract_snippets", d_x_wc, ldx, C, W, nt, twav_min, loc_range, long_range,
                                     d_clips, max_clips, n_found))
        return st;
    hipStream_t s = (hipStream_t)hip_stream;
    const SnippetScratch z(C, W);
    Scratch sc(s);
    unsigned *pk, *iso;
    int* bsum;
    long long* total;
    sc.add(&pk, z.nwords);
    sc.add(&iso, z.nwords);
    sc.add(&bsum, z.nb);
    sc.add(&total, 1);
    TRY_HIP(sc.commit());
    return snippet_run(d_x_wc, ldx, C, W, nt, twav_min, th_single_ch, loc_range, long_range, d_clips, max_clips,
                       d_chan, d_time, n_found, z, pk, iso, bsum, total, s);
}

lss_status lss_extract_snippets_noalloc(const float* d_x_wc, int64_t ldx, int32_t C, int64_t W, int32_t nt,
                                        int32_t twav_min, float th_single_ch, const int32_t* loc_range,
                                        const int32_t* long_range, float* d_clips, int64_t max_clips, int32_t* d_chan,
                                        int32_t* d_time, int64_t* n_found, void* hip_stream) {
    if (lss_status st = snippet_args("lss_extract_snippets_noalloc", d_x_wc, ldx, C, W, nt, twav_min, loc_range,
                                     long_range, d_clips, max_clips, n_found))
        return st;
    const SnippetScratch z(C, W);
    const size_t need = 2 * round256(sizeof(unsigned) * z.nwords) + round256(sizeof(int) * z.nb) + 256;
    if (need > sizeof(scratch))
        return err(LSS_ERR_ARG, "lss_extract_snippets_noalloc: C = " + std::to_string(C) + ", W = " +
                                    std::to_string(W) + " needs " + std::to_string(need) +
                                    " bytes of scratch, more than the fixed " + std::to_string(sizeof(scratch)) +
                                    " (C * ceil(W / 64) must stay below about 520000)");
    int dev = 0;
    TRY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) return err(LSS_ERR_ARG, "lss_extract_snippets_noalloc: device index too large");
    hipStream_t s = (hipStream_t)hip_stream;
    std::lock_guard<std::mutex> g(g_noalloc_mu[dev]);
    char* base = nullptr;
    if (lss_status st = noalloc_scratch(dev, &base)) return st;
    // the kernels are done before the lock is released, on every path
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{s};
    unsigned* pk = (unsigned*)base;
    unsigned* iso = (unsigned*)(base + round256(sizeof(unsigned) * z.nwords));
    int* bsum = (int*)(base + 2 * round256(sizeof(unsigned) * z.nwords));
    long long* total = (long long*)(base + 2 * round256(sizeof(unsigned) * z.nwords) + round256(sizeof(int) * z.nb));
    return snippet_run(d_x_wc, ldx, C, W, nt, twav_min, th_single_ch, loc_range, long_range, d_clips, max_clips,
                       d_chan, d_time, n_found, z, pk, iso, bsum, total, s);
}

lss_status lss_wpca(const float* d_clips, int64_t n, int32_t nt, int32_t n_pcs, double* d_wpca, double* d_sv,
                    void* hip_stream) {
    if (!d_clips || !d_wpca || !d_sv || n < 1 || nt < 1 || nt > lss::kEigMax || n_pcs < 1 || n_pcs > nt)
        return err(LSS_ERR_ARG, "lss_wpca: bad argument (nt <= 96, 1 <= n_pcs <= nt)");
    hipStream_t s = (hipStream_t)hip_stream;
    int S;
    int64_t kchunk;
    lss::gemm64_split(nt, nt, n, 1, &S, &kchunk);
    Scratch sc(s);
    double *Q, *P, *G;
    sc.add(&Q, (size_t)n * nt);
    sc.add(&P, (size_t)S * nt * nt);
    sc.add(&G, (size_t)nt * nt);
    TRY_HIP(sc.commit());
    TRY_HIP(lss::launch_normalize_clips(d_clips, n, nt, Q, s));
    TRY_HIP(lss::launch_gemm64_nk(Q, n, Q, n, P, nt, nt, n, S, kchunk, 1, s));
    TRY_HIP(lss::launch_reduce64(P, S, nt, nt, 1, 1.0, 0.0, 0.0, G, nt, s));
    TRY_HIP(lss::launch_pca(G, nt, nt, n_pcs, d_wpca, d_sv, s));
    TRY_HIP(hipStreamSynchronize(s));
    return LSS_OK;
}

lss_status lss_train_release(void) {
    std::lock_guard<std::mutex> g(g_arena_mu);
    for (auto& kv : g_arena) {
        if (!kv.second.p) continue;
        int cur = 0;
        TRY_HIP(hipGetDevice(&cur));
        TRY_HIP(hipSetDevice(kv.first));
        TRY_HIP(hipDeviceSynchronize());
        TRY_HIP(hipFree(kv.second.p));
        TRY_HIP(hipSetDevice(cur));
        kv.second = Arena{};
    }
    return LSS_OK;
}

}  // extern "C"
