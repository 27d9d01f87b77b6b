// Launchers of the gfx950 kernels of liblss.  All take a hipStream_t and return hipError_t of the
// launch.  Layouts: [a][b] is row-major with the last index fastest; "ld" is the row stride in
// elements.  X_cw = [C][ldw] channel-major window, X_wc = [W][C] time-major window.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lss {

// ---- preprocessing (preprocess.hip) ---------------------------------------------------------
// Exact int64 channel sums of an int16 [W][C] window.
// Exact int64 channel sums: per-block partials part [channel_sum_blocks(W)][C], then sums[C].
int64_t channel_sum_blocks(int64_t W);
hipError_t launch_channel_sums(const int16_t* win, int64_t W, int C, long long* part, long long* sums,
                               hipStream_t s);
// out[c][w] = (x[w][c] - mean_c) - lower_median_c'(x[w][c'] - mean_c'), transposed to [C][ldo].
// do_car = 0 skips the median.  C <= 1024.
hipError_t launch_car_transpose(const int16_t* win, int64_t W, int C, const long long* sums,
                                int do_car, float* out, int64_t ldo, hipStream_t s);

// Window assembly from device samples: segment k writes n int16 at dst, from src (period 0:
// contiguous) or repeating the period-element row src; up to kGatherMaxSeg segments per launch
// (launch_gather splits longer lists).  chunk0 is set by the launcher.
struct GatherSeg {
    int16_t* dst;
    const int16_t* src;
    int64_t n;
    int64_t period;
    int64_t chunk0;
};
constexpr int kGatherMaxSeg = 32;
struct GatherArgs {
    GatherSeg seg[kGatherMaxSeg];
    int n;
    int64_t chunks;
};
hipError_t launch_gather(const GatherSeg* segs, int n, hipStream_t s);

// Overlap-save highpass: y[c][i] = sum_{j=-Llo..Lhi} g[j] x[c][(i+j) mod W].
constexpr int kFftN = 8192;          // FFT length of one segment
constexpr int kFftLog2 = 13;
constexpr int kHpHalf = 800;         // taps kept on each side of the filter centre
// spec: [kFftN] complex (float2) spectrum of the time-reversed kernel, scaled by 1/N;
// tw: [kFftN] float2 twiddles exp(-2 pi i m / N).
hipError_t launch_highpass(const float* x, int64_t ldx, float* y, int64_t ldy, int C, int64_t W,
                           int Llo, int Lhi, const float2* spec, const float2* tw, hipStream_t s);
// spec for window length W from the device taps [ft] (FP64 sums, FP32 result); tw64: [kFftN]
// exp(-2 pi i m / N) in FP64.  Llo, Lhi = min(kHpHalf, W/2), min(kHpHalf, W-1-W/2).
hipError_t launch_hp_spectrum(const float* taps, int ft, int64_t W, int Llo, int Lhi, const double2* tw64,
                              float2* spec, hipStream_t s);
// Plain batched forward FFT of [nb][kFftN] complex rows (test hook of the LDS FFT).
hipError_t launch_fft_rows(float2* data, int nb, const float2* tw, hipStream_t s);

// ---- contractions (gemm.hip, project.hip) ---------------------------------------------------
// C[m][n] = sum_k A(m,k) B(n,k);  A(m,k) = A[m*lda+k] (a_km=0) or A[k*lda+m] (a_km=1);
// B(n,k) = B[n*ldb+k].  FP32 MFMA, FP32 accumulation.  lda, ldb multiples of 4; 16-byte aligned.
// tile: output tile shape 0 = 128x128, 1 = 128x64, 2 = 64x128, 3 = 64x64 (results identical).
constexpr int kGemmTiles = 4;
hipError_t launch_gemm_f32(const float* A, int64_t lda, int a_km, const float* B, int64_t ldb,
                           float* C, int64_t ldc, int M, int N, int K, hipStream_t s, int tile = 0);
// Precision variants with operands converted on the fly (prec = lss_precision).
hipError_t launch_gemm_lp(int prec, const float* A, int64_t lda, int a_km, const float* B,
                          int64_t ldb, float* C, int64_t ldc, int M, int N, int K, hipStream_t s, int tile = 0);
// Xp[w][k*C+c] = sum_m X[w+m-nt/2][c] * wpca[k][m] (zero outside [0,W)); X rows have stride ldx,
// Xp rows ldp.  K <= 8, nt <= 127.
// tile: workgroup shape 0..3 (0 = by window length; results identical for every shape).
hipError_t launch_project(const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int K,
                          int nt, float* xp, int64_t ldp, hipStream_t s, int tile = 0);
// The same product on the matrix cores with split operands (prec = LSS_PREC_BF16X3 / FP16X3).
hipError_t launch_project_x3(int prec, const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int K,
                             int nt, float* xp, int64_t ldp, hipStream_t s, int tile = 0);

// ---- pursuit (pursuit.hip) ------------------------------------------------------------------
// Row strides ldT (B, ctc, nm) and ldC (X, waves) are multiples of 4 (float4 access); padding
// columns are never read as scores.
struct PursuitArgs {
    float* B;              // [W][ldT] scores (updated)
    float* X;              // [W][ldC] residual (updated by the final residual kernel)
    const float* nm;       // [ldT]
    const float* ctc;      // [T][2nt+1][ldT]  ctc[j][l][i]
    const float* waves;    // [T][nt][ldC]
    float* cfmax;          // [W]
    int* imax;             // [W]
    unsigned char* flags;  // peak flag words of the current pass (uint32 [ceil(W/32)])
    int* sp_col;           // [cap] peak column
    int* sp_tmp;           // [cap] template
    float* sp_amp;         // [cap]
    int* pass_off;         // [max_passes + 1] first spike of each pass that found spikes
    float* amprow;         // persistent pursuit: the Cfmax/imax row's amplitudes (set by the launcher)
    // measurement mode: the update kernel's workgroup b writes its GPU-clock (100 MHz) start and end
    // to ts[2b], ts[2b + 1] (NULL otherwise)
    unsigned long long* ts;
    // persistent pursuit: Cfmax, imax and amplitude rows [2][kPersistMaxW], one per pass parity,
    // and the workgroups' barrier flags [kPersistMaxGrid][kBarStride]
    float* prow_cf;
    int* prow_im;
    float* prow_amp;
    int* bar;
    // [0] spikes of committed passes, [1] spikes of the last pass, [2] capacity overflow,
    // [3] converged (a pass found nothing), [4] passes that found spikes, [5] residual list overflow (one pass),
    // [7] grid-barrier timeout of the persistent pursuit, [8] the
    // peak kernel's workgroup ticket, [kEpoch] the persistent pursuit's barrier epoch (advanced, not
    // zeroed, by launch_reset_counters) (kCounters ints)
    int* counters;
    int64_t W;
    int64_t ldT, ldC;
    int T, C, K, nt;
    int cap;
    float th2;
};
constexpr int kCounters = 16;
constexpr int kEpoch = 15;
// counters[0..kCounters) = 0 except counters[kEpoch], incremented (bar: the persistent pursuit's
// barrier flags [kPersistMaxGrid][kBarStride], raised to the epoch's base every 2^16 epochs; may be
// null)
hipError_t launch_reset_counters(int* counters, hipStream_t s, int* bar = nullptr);
hipError_t launch_score_all(const PursuitArgs& a, hipStream_t s);     // T <= 2048
hipError_t launch_peaks(const PursuitArgs& a, hipStream_t s);
// grid = workgroups (grid-stride over the pass's work items; any grid gives the same result);
// start/stop (measurement mode): events recorded around the launch; cpw = score columns per
// wavefront, 1 or 2 (same results)
hipError_t launch_update(const PursuitArgs& a, int grid, hipStream_t s, hipEvent_t start = nullptr,
                         hipEvent_t stop = nullptr, int cpw = 2);
// measurement mode: out[2p], out[2p + 1] = min start, max end of the update stamps ts[p][grid][2]
hipError_t launch_stamp_span(const unsigned long long* ts, int launches, int grid, unsigned long long* out,
                             hipStream_t s);
// Persistent pursuit of windows W <= kPersistMaxW: passes [first_pass, max_passes) in one launch of
// co-resident workgroups (persistent_grid of `grid`: hipErrorInvalidConfiguration when that is 0;
// the pursuit then belongs on the per-pass kernels); with first_pass = 0 it also does
// the first scoring, otherwise it continues from the state that first_pass per-pass (peaks,
// update) rounds left.  Results equal those of score_all + max_passes x (peaks, update).
// counters[7] = 1 reports a barrier timeout.
constexpr int kPersistMaxW = 4096;
constexpr int kPersistMaxGrid = 2048;   // workgroups of one persistent launch (barrier flags)
constexpr int kBarStride = 16;          // ints between two barrier flags (one per 64 bytes)
// used (may be NULL): the workgroups launched.  With a.ts set (measurement mode) workgroup b writes
// its GPU-clock start and end to ts[2b], ts[2b + 1].
hipError_t launch_pursuit_persistent(const PursuitArgs& a, int grid, int first_pass, int max_passes,
                                     hipStream_t s, int* used = nullptr);
// Workgroups of the persistent pursuit that the device holds at once for T templates, window W (0
// when T > 1024: those pursuits run on the per-pass kernels).  Each workgroup keeps the score rows
// of ceil(W / grid) columns in its LDS, so a grid also has a minimum (persistent_grid).
hipError_t persistent_max_grid(int T, int64_t W, int* grid);
// The grid a launch of `a` at `want` workgroups runs: the largest one <= want whose workgroups, with
// the LDS it needs, the device holds at once; 0 when the window's score rows do not fit the LDS of
// any such grid (the pursuit then runs on the per-pass kernels).
hipError_t persistent_grid(const PursuitArgs& a, int want, int* grid);
bool persistent_supports(int T);
// residual of all passes in pass order (even, then odd spikes of a pass), after the last pass
hipError_t launch_residual(const PursuitArgs& a, hipStream_t s);

struct FeatureArgs {
    const float* X;      // [W][ldC] final residual
    const float* wpca;   // [K][nt]
    const int* iCC;      // [nC][C]
    const int* iU;       // [T]
    const float* Ucc;    // [nC][T][K]
    const float* xc;     // [C]
    const float* yc;     // [C]
    const int* sp_col;
    const int* sp_tmp;
    const float* sp_amp;
    const int* counters; // spike count = counters[0] + counters[1] (or n when NULL)
    float* tF;           // [n][nC][K]
    float* xs;           // [n]
    float* ys;           // [n]
    int n, cap, C, T, K, nt, nC;
    int64_t W, ldC;
};
hipError_t launch_features(const FeatureArgs& a, int grid, hipStream_t s);

// Counters and the first min(counters[0] + counters[1], cap) spikes into host-mapped (pinned)
// buffers: h_counters[8], h_col/h_tmp/h_amp/h_xs/h_ys [cap], h_tF [cap][FK].
struct StageArgs {
    const int* counters;
    const int* col;
    const int* tmp;
    const float* amp;
    const float* tF;
    const float* xs;
    const float* ys;
    int* h_counters;
    int* h_col;
    int* h_tmp;
    float* h_amp;
    float* h_tF;
    float* h_xs;
    float* h_ys;
    int cap, FK;
};
hipError_t launch_stage_spikes(const StageArgs& a, hipStream_t s);

// ---- grouped launches: the pursuit tail of several windows (one per handle) in one launch each ----
// Member i of a group takes workgroups [first[i], first[i + 1]) of a launch and runs on them what
// the single-window kernel would run on a grid of its own, so results equal those launches'.
constexpr int kGroupMax = 8;
constexpr int kGroupFeatureGrid = 256;   // workgroups per member of a grouped features launch
struct GroupMember {
    PursuitArgs a;   // the window's pursuit (counters, bar: its counter reset)
    FeatureArgs f;   // its features (counters set)
    StageArgs s;     // its readback
};
// d: the members in device memory (read by the kernels), m: the same members on the host (their
// sizes shape the launch); n <= kGroupMax.
hipError_t launch_group_reset(const GroupMember* d, int n, hipStream_t s);
// Persistent pursuits from pass 0: member i on grids[i] workgroups (a grid that persistent_grid
// gives for it) up to max_passes[i]; every member's workgroups must be co-resident together
// (group_persistent_fits), which the LDS of the largest member per workgroup decides.
hipError_t group_persistent_fits(const GroupMember* m, const int* grids, int n, bool* fits);
hipError_t launch_group_persistent(const GroupMember* d, const GroupMember* m, const int* grids,
                                   const int* max_passes, int n, hipStream_t s);
hipError_t launch_group_residual(const GroupMember* d, const GroupMember* m, int n, hipStream_t s);
// Features of every member, each also written with its spikes and counters to the member's
// readback staging (no separate staging launch).
hipError_t launch_group_features(const GroupMember* d, const GroupMember* m, int n, hipStream_t s);

// ---- template refinement (adjoint.hip) --------------------------------------------------------
// Residual energy of a fixed spike set (sorted by column) and its gradient in U and wPCA.
struct RefineArgs {
    const float* X;       // [W][ldx] whitened window
    int64_t ldx;
    int vec_x;            // X rows may be read as float4 (ldx % 4 == 0, 16-byte aligned)
    const int* col;       // [n] spike columns, ascending
    const int* tmp;       // [n] templates
    const float* amp;     // [n]
    int n;
    const float* U;       // [T][K][C]
    const float* wpca;    // [K][nt]
    const float* waves;   // [T][nt][ldC] from U, wpca (launch_waves)
    float* R;             // [W][ldC] residual (written by launch_refine_residual)
    double* part;         // [refine_blocks(W)] partial sums of R^2
    double* loss;         // [1]
    float* gU;            // [T][K][C]
    float* dwpart;        // [T][K][nt]
    float* dupart;        // [T][refine_chunks(nt)][K][C] per-row-chunk shares of dU
    float* gW;            // [K][nt]
    int* flag;            // [1] set when a template id is out of range
    int64_t W, ldC;
    int T, C, K, nt;
};
int refine_blocks(int64_t W);
int refine_chunks(int nt);
hipError_t launch_fill_i32(int* p, int n, int v, hipStream_t s);   // p[0..n) = v (capturable)
// out = x - alpha * y elementwise (FP64 arithmetic, FP32 result)
hipError_t launch_step(float* out, const float* x, const float* y, double alpha, int64_t n, hipStream_t s);
// *out = sum of squares of x in FP64 (fixed order); part: 256 doubles of scratch
hipError_t launch_sumsq(const float* x, int64_t n, double* part, double* out, hipStream_t s);
// R and loss = sum R^2 (C <= 1024, nt <= 127)
hipError_t launch_refine_residual(const RefineArgs& a, hipStream_t s);
// gU, gW from R (after launch_refine_residual; K <= 8)
hipError_t launch_refine_adjoint(const RefineArgs& a, hipStream_t s);

// ---- load-time tables in FP64 (tables.hip) --------------------------------------------------
// nm[t] = sum U^2; waves[t][m][c] = sum_k U[t][k][c] wpca[k][m]; ctc[j][l][i] (see lss.h).
// U rows (one per template) have stride ldu >= K*C; waves rows ldC, ctc rows ldT.
// waves[t][m][c] = sum_k U[t][k][c] wpca[k][m] (FP64 accumulation), rows of stride ldC
hipError_t launch_waves(const float* U, int64_t ldu, const float* wpca, int T, int K, int C, int nt,
                        float* waves, int64_t ldC, hipStream_t s);
// ws: device workspace of tables_workspace_doubles(T, K, C, nt) doubles (the Gram matrix, WtW
// and the Gram GEMM's split-K partials).
size_t tables_workspace_doubles(int T, int K, int C, int nt);
// iU and Ucc of prepare_extract from the device templates (one workgroup per template).
hipError_t launch_extract(const float* U, int64_t ldu, int T, int K, int C, const int* iCC, int nC, int* iU,
                          float* Ucc, hipStream_t s);
hipError_t launch_tables(const float* U, int64_t ldu, const float* wpca, int T, int K, int C, int nt,
                         float* nm, float* waves, int64_t ldC, float* ctc, int64_t ldT, double* ws, hipStream_t s);

// ---- training-side artifacts in FP64 (train.hip) -------------------------------------------------
// P[z][m][n] = sum over K-chunk z of A(m,k) B(k,n), A(m,k) = A[m*lda+k]; "_nk": B(k,n) = B[n*ldb+k],
// "_kn": B(k,n) = B[k*ldb+n].  sym (A = B, M = N) skips tiles below the diagonal.  gemm64_split
// picks S chunks of kchunk (a multiple of 32) for about 1024 workgroups; P holds S*M*N doubles.
void gemm64_split(int M, int N, int64_t K, int sym, int* S, int64_t* kchunk);
hipError_t launch_gemm64_f32_nk(const float* A, int64_t lda, const float* B, int64_t ldb, double* P, int M, int N,
                                int64_t K, int S, int64_t kchunk, int sym, hipStream_t s);
hipError_t launch_gemm64_nk(const double* A, int64_t lda, const double* B, int64_t ldb, double* P, int M, int N,
                            int64_t K, int S, int64_t kchunk, int sym, hipStream_t s);
hipError_t launch_gemm64_kn(const double* A, int64_t lda, const double* B, int64_t ldb, double* P, int M, int N,
                            int64_t K, int S, int64_t kchunk, hipStream_t s);
// out = beta*out + alpha*sum_z P[z] (+ diag on the diagonal), chunks summed in order
hipError_t launch_reduce64(const double* P, int S, int M, int N, int sym, double alpha, double beta, double diag,
                           double* out, int64_t ldo, hipStream_t s);
hipError_t launch_sumsq64(const double* x, int M, int N, int64_t ldx, double* out, hipStream_t s);
hipError_t launch_f64_to_f32(const double* x, int M, int N, int64_t ldx, float* y, int64_t ldy, hipStream_t s);
// K[i][j] = exp(-((xc_i-xc_j)^2 + (yc_i-shift_i-yc_j)^2) / (2 sig^2)) + lam [i==j]; shift may be NULL
hipError_t launch_interp_kernel(const double* xc, const double* yc, const double* shift, int C, double sig, double lam,
                                double* Kout, hipStream_t s);
// Jacobi eigensolver in LDS, matrices of order n <= kEigMax
constexpr int kEigMax = 96;
// row j of W = first row of the ZCA of cc restricted to idx[j][0..nr), scattered to columns idx[j]
hipError_t launch_whiten_local(const double* cc, int64_t ldcc, const int* idx, int nr, int C, double eps, double* W,
                               int64_t ldw, hipStream_t s);
hipError_t launch_whiten_full(const double* cc, int64_t ldcc, int n, double eps, double* W, int64_t ldw, hipStream_t s);
// out [npc][n]: leading eigenvectors of G (sign: largest-magnitude entry positive); sv [npc]
hipError_t launch_pca(const double* G, int64_t ldg, int n, int npc, double* out, double* sv, hipStream_t s);
// Isolated peaks of a [W][ldx] window: bitmaps pk, iso of C rows x nw words (nw >= ceil(W/32)).
hipError_t launch_isolated_peaks(const float* X, int64_t ldx, int C, int64_t W, int lc, int lt, float th, int gc,
                                 int gt, int edge, unsigned* pk, unsigned* iso, int64_t nw, hipStream_t s);
int snippet_scan_blocks(int64_t nwords);
// clips[i][m] = X[t_i - twav_min + m][c_i] for the iso bits in (channel, time) order, i < max_clips;
// bsum: snippet_scan_blocks(nwords) ints; *total = number of iso bits
hipError_t launch_emit_snippets(const unsigned* iso, int64_t nwords, int64_t nw, int* bsum, long long* total,
                                const float* X, int64_t ldx, int nt, int twav_min, float* clips, int64_t max_clips,
                                int* chan, int* tim, hipStream_t s);
// Q[m][i] = clips[i][m] / ||clips[i]|| (FP64)
hipError_t launch_normalize_clips(const float* clips, int64_t n, int nt, double* Q, hipStream_t s);

// ---- feature-space cluster assignment (assign.hip) ----
hipError_t launch_assign_clusters(const float* tF, int64_t n, int D, const float* mu, int K, const float* xs,
                                  const float* ys, const float* cx, const float* cy, float R2, int* label,
                                  hipStream_t s);
hipError_t launch_cluster_centroids(const float* Ucc, int T, int nC, int K, float* out, hipStream_t s);

}  // namespace lss
