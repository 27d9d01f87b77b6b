// liblss C ABI: sorter handle, artifact upload, per-window pipeline.
//
// The handle plays the role of OnlineSpikesV2 (src/SpikeSorter/OnlineSpikesV2.cu:247-630: load
// artifacts, allocate every buffer once for the largest window) and of its per-batch operator
// kilosortMatchingPursuit + computeClosestClusters (OnlineSpikesV2.cu:823-971), with Kilosort4
// semantics where the two differ.  All device work runs on the handle's stream; the pursuit runs
// its max_passes passes without host round trips (a pass that finds no peak ends the work of the
// later ones on the device) and the counters are read back once per window.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/lss.h"
#include "lss_kernels.h"
#include "npy.h"

namespace {

thread_local std::string g_err;

lss_status fail(lss_status st, const std::string& msg) {
    g_err = msg;
    return st;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(LSS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        n = count;
        return hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(count, 1));
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct Handle {
    lss_params prm{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int C = 0, T = 0, K = 0, nt = 0, nC = 0;
    int64_t Wmax = 0, ldw = 0;
    // row strides, multiples of 4 (float4 access): U/Xp rows (K*C), whitening matrix, X, waves (C),
    // B, ctc, nm (T)
    int ldq = 0, ldm = 0;
    int64_t ldC = 0, ldT = 0;

    // tables
    DevBuf<float> U, wpca, wmat, nm, waves, ctc, Ucc, xc, yc;
    DevBuf<double> tables_ws;   // Gram matrix + WtW of the FP64 table build (kept for refinement)
    DevBuf<double> wrot64;      // [C][C] Wrot in FP64 (per-batch drift: wmat = M @ Wrot)
    DevBuf<double> drift_ws;    // split-K partials and the FP64 product of lss_set_drift
    DevBuf<int> iCC, iU;
    DevBuf<float2> tw;
    DevBuf<float> hp_dev;        // highpass taps (hp_filter) on the device
    DevBuf<double2> tw64;        // exp(-2 pi i m / N) in FP64, for the spectra
    // overlap-save spectra per window length: built on the device on first use, kept; entries a
    // streaming batcher captured in its graphs (pins > 0) are never evicted
    struct Spec {
        float2* p = nullptr;
        int lo = 0, hi = 0;
        int pins = 0;
        uint64_t used = 0;
    };
    std::map<int64_t, Spec> specs;
    uint64_t spec_clock = 0;
    std::vector<float> h_xc, h_yc;
    int64_t last_W = 0;          // window length of the last pursuit (template refinement reads it)

    // window buffers
    DevBuf<int16_t> win;
    DevBuf<long long> sums, sum_part;   // channel sums and their per-block partials
    DevBuf<float> xcw, xhp, xw, xp, B, cfmax;
    // the streaming batcher's second set (allocated by its first step that runs ahead): the next
    // batch's window, whitened data and scores, preprocessed while the current batch's pursuit runs,
    // then swapped with win / xw / B when that batch's step starts
    DevBuf<int16_t> win_alt;
    DevBuf<float> xw_alt, B_alt;
    uint64_t pre_version = 0;    // advanced when an input of the preprocessing changes (drift, templates)
    // a stream's pursuit launched ahead of its step (pipelined steps) that uses this handle: every
    // entry point that takes the handle first waits for it and marks the handle touched, which makes
    // that step discard the launch and sort its batch again
    hipEvent_t pending = nullptr;
    bool touched = false;
    DevBuf<int> imax;
    DevBuf<unsigned char> flags;
    DevBuf<int> sp_col, sp_tmp, counters, pass_off;
    DevBuf<float> prow_cf, prow_amp;   // persistent pursuit rows [2][kPersistMaxW] (pass parity)
    DevBuf<int> prow_im;
    DevBuf<int> pbar;                   // persistent pursuit barrier flags [kPersistMaxGrid][kBarStride]
    DevBuf<float> sp_amp, tF, xs, ys;
    int* h_counters = nullptr;   // pinned
    // pinned readback staging of the batched/streaming path (first stage_cap spikes)
    int stage_cap = 0;
    int* hs_counters = nullptr;
    int *hs_col = nullptr, *hs_tmp = nullptr;
    float *hs_amp = nullptr, *hs_tF = nullptr, *hs_xs = nullptr, *hs_ys = nullptr;
    lss::StageArgs stage{};      // device-side view of the staging buffers

    // template refinement (allocated on first use): residual, waves of the given U/wPCA, partials
    DevBuf<float> rf_R, rf_R2, rf_waves, rf_dwpart, rf_dupart, rf_U, rf_U1, rf_gU, rf_gW, rf_W1, rf_amp;
    DevBuf<double> rf_sums;
    DevBuf<int> rf_col, rf_tmp;
    DevBuf<double> rf_part, rf_loss;
    DevBuf<int> rf_flag;
    hipEvent_t rf_ev[3] = {};
    float rf_ms[3] = {};         // residual (with waves), adjoint, total of the last call
    // the streaming batcher's refinement in the background: its stream, a copy of the window's
    // residual and spike lists, and the second template set it writes (copied over the handle's
    // own when the batcher adopts it)
    hipStream_t rf_stream = nullptr;
    std::atomic<bool> rf_busy{false};   // a background refinement holds the refinement buffers
    DevBuf<float> sn_x, sn_amp;
    DevBuf<int> sn_col, sn_tmp;
    DevBuf<float> sh_U, sh_wpca, sh_nm, sh_waves, sh_ctc, sh_Ucc;
    DevBuf<int> sh_iU;

    hipEvent_t ev[8] = {};
    // per-launch timing of the pursuit update kernel (measurement mode)
    bool prof_update = false;
    int upd_grid = 8192;        // workgroups of the update kernel (autotuned)
    int upd_cpw = 2;            // score columns per wavefront of the update kernel (autotuned)
    int tuned_passes = 0;       // passes in the first graph segment of the batcher (0 = all)
    int pers_grid = 128;        // workgroups of the persistent pursuit (windows <= kPersistMaxW; 0 = off)
    int pers_first = 0;         // leading passes on the per-pass kernels before the persistent kernel
    int tile_whiten = 0;        // GEMM output tile shapes (launch_gemm_*: 0 = 128x128 .. 3 = 64x64), autotuned
    int tile_scores = 0;
    int proj_tile = 0;          // wPCA projection workgroup shape (launch_project*), autotuned
    std::vector<hipEvent_t> upd_ev;
    float upd_ms = 0.f;
    int upd_launches = 0;
    std::vector<float> upd_each;   // per-launch ms of the last window (measurement mode): HIP events
    std::vector<float> upd_gpu;    // the same from the workgroups' GPU-clock stamps
    std::vector<int> upd_spikes;   // spikes each of those launches applied
    DevBuf<unsigned long long> upd_ts;   // [max_passes][upd_grid][2] GPU-clock stamps
    DevBuf<unsigned long long> upd_span; // [max_passes][2] first start, last end per launch
    // measurement mode, persistent pursuit: its workgroups' stamps, the span, its first pass and
    // grid in the last window (pers_p0 < 0: it did not run) and what it did
    DevBuf<unsigned long long> pers_ts, pers_span;
    int pers_p0 = -1, pers_used = 0;
    float pers_ms = 0.f;
    int pers_spikes = 0, pers_passes = 0;
    float last_ms[8] = {};   // car, highpass, whiten, project, scores GEMM, pursuit, features, total
    lss_load_info load{};    // which extraction tables were given, the ctc.npy comparison
    bool alive = false;

    ~Handle() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : upd_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : rf_ev)
            if (e) (void)hipEventDestroy(e);
        if (h_counters) (void)hipHostFree(h_counters);
        for (void* q : {(void*)hs_counters, (void*)hs_col, (void*)hs_tmp, (void*)hs_amp, (void*)hs_tF,
                        (void*)hs_xs, (void*)hs_ys})
            if (q) (void)hipHostFree(q);
        for (auto& kv : specs)
            if (kv.second.p) (void)hipFree(kv.second.p);
        if (rf_stream) (void)hipStreamDestroy(rf_stream);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

lss_status check_params(const lss_params& p) {
    if (p.n_chan < 1 || p.n_chan > 1024) return fail(LSS_ERR_ARG, "n_chan must be in [1, 1024]");
    if (p.n_templates < 1 || p.n_templates > 2048) return fail(LSS_ERR_ARG, "n_templates must be in [1, 2048]");
    if (p.n_pcs < 1 || p.n_pcs > 8) return fail(LSS_ERR_ARG, "n_pcs must be in [1, 8]");
    if (p.nt < 3 || p.nt > 127 || (p.nt % 2) == 0) return fail(LSS_ERR_ARG, "nt must be odd in [3, 127]");
    if (p.nt0min < 0 || p.nt0min >= p.nt) return fail(LSS_ERR_ARG, "nt0min must be in [0, nt)");
    if (p.n_nearest < 1 || p.n_nearest > p.n_chan || p.n_nearest * p.n_pcs > 256)
        return fail(LSS_ERR_ARG, "n_nearest must be in [1, n_chan] with n_nearest*n_pcs <= 256");
    if (p.max_window < 2 * p.nt + 2) return fail(LSS_ERR_ARG, "max_window too small");
    if (p.max_passes < 1 || p.max_passes > 256) return fail(LSS_ERR_ARG, "max_passes must be in [1, 256]");
    if (p.max_spikes < 1) return fail(LSS_ERR_ARG, "max_spikes must be >= 1");
    if (p.precision < LSS_PREC_FP32 || p.precision > LSS_PREC_FP16X3) return fail(LSS_ERR_ARG, "bad precision");
    if (p.dup_bins < 0) return fail(LSS_ERR_ARG, "dup_bins must be >= 0");
    return LSS_OK;
}

// Stable argsort of squared distances (prepare_extract, template_matching.py:20-26; ties by index)
void nearest_channels(const std::vector<float>& xc, const std::vector<float>& yc, int nC, std::vector<int>& iCC) {
    const int C = (int)xc.size();
    iCC.assign((size_t)nC * C, 0);
    std::vector<int> idx(C);
    std::vector<double> ds(C);
    for (int c = 0; c < C; ++c) {
        for (int j = 0; j < C; ++j) {
            // ds[j] in the reference's float64 arithmetic: (xc - xc[c])^2 + (yc - yc[c])^2
            double dx = (double)xc[j] - (double)xc[c], dy = (double)yc[j] - (double)yc[c];
            ds[j] = dx * dx + dy * dy;
        }
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ds[a] < ds[b]; });
        for (int j = 0; j < nC; ++j) iCC[(size_t)j * C + c] = idx[j];
    }
}

// Spectrum of the overlap-save kernel for window length W (the fft_highpass pad/crop rule,
// kilosort/preprocessing.py:138-154, then fftshift), from the handle's cache or built on the
// device in FP64 on the handle's stream (ordered before any highpass that reads it).  pin = a
// streaming batcher captures the buffer in its graphs: the entry then stays until unpinned.
constexpr size_t kSpecCache = 16;

lss_status get_spectrum(Handle& h, int64_t W, bool pin, const Handle::Spec** out) {
    auto it = h.specs.find(W);
    if (it == h.specs.end()) {
        if (h.specs.size() >= kSpecCache) {
            auto victim = h.specs.end();
            for (auto j = h.specs.begin(); j != h.specs.end(); ++j)
                if (j->second.pins == 0 && (victim == h.specs.end() || j->second.used < victim->second.used))
                    victim = j;
            if (victim != h.specs.end()) {
                HIP_TRY(hipStreamSynchronize(h.stream));   // no queued highpass still reads it
                HIP_TRY(hipFree(victim->second.p));
                h.specs.erase(victim);
            }
        }
        Handle::Spec sp;
        sp.lo = (int)std::min<int64_t>(lss::kHpHalf, W / 2);
        sp.hi = (int)std::min<int64_t>(lss::kHpHalf, W - 1 - W / 2);
        HIP_TRY(hipMalloc((void**)&sp.p, sizeof(float2) * lss::kFftN));
        hipError_t e = lss::launch_hp_spectrum(h.hp_dev.p, (int)h.hp_dev.n, W, sp.lo, sp.hi, h.tw64.p, sp.p, h.stream);
        if (e != hipSuccess) {
            (void)hipFree(sp.p);
            return fail(LSS_ERR_HIP, std::string("launch_hp_spectrum: ") + hipGetErrorString(e));
        }
        it = h.specs.emplace(W, sp).first;
    }
    it->second.used = ++h.spec_clock;
    if (pin) ++it->second.pins;
    *out = &it->second;
    return LSS_OK;
}

// iU (channel of largest template energy) and Ucc = U[:, :, iCC[:, iU]] for host U [T][K][C]
// (prepare_extract, template_matching.py:20-26), uploaded to the handle.
// iU / Ucc (prepare_extract, template_matching.py:20-26) from the device templates; iCC is on the
// device already.  Enqueued on the handle's stream.
lss_status refresh_extract(Handle& h) {
    HIP_TRY(lss::launch_extract(h.U.p, h.ldq, h.T, h.K, h.C, h.iCC.p, h.nC, h.iU.p, h.Ucc.p, h.stream));
    return LSS_OK;
}

lss_status create_impl(const lss_params* p, const lss_artifacts* a, Handle** out) {
    if (!p || !a || !out) return fail(LSS_ERR_ARG, "null argument");
    if (lss_status st = check_params(*p)) return st;
    if (!a->U || !a->wPCA || !a->Wrot || !a->hp_filter || a->hp_len < 1 || !a->xc || !a->yc)
        return fail(LSS_ERR_ARG, "missing artifact array");
    std::unique_ptr<Handle> h(new Handle());
    h->prm = *p;
    HIP_TRY(hipGetDevice(&h->device));
    HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->own_stream = true;
    const int C = p->n_chan, T = p->n_templates, K = p->n_pcs, nt = p->nt, nC = p->n_nearest;
    h->C = C; h->T = T; h->K = K; h->nt = nt; h->nC = nC;
    h->Wmax = p->max_window;
    h->ldw = (h->Wmax + 63) / 64 * 64;
    h->ldq = (K * C + 3) / 4 * 4;
    h->ldm = (C + 3) / 4 * 4;
    h->ldC = h->ldm;
    h->ldT = (T + 3) / 4 * 4;
    const int L = 2 * nt + 1;
    const int64_t Wm = h->Wmax;

    // ---- tables ----
    HIP_TRY(h->U.alloc((size_t)T * h->ldq));
    HIP_TRY(h->wpca.alloc((size_t)K * nt));
    HIP_TRY(h->wmat.alloc((size_t)C * h->ldm));
    HIP_TRY(h->nm.alloc(h->ldT));
    HIP_TRY(h->waves.alloc((size_t)T * nt * h->ldC));
    HIP_TRY(h->ctc.alloc((size_t)T * L * h->ldT));
    HIP_TRY(hipMemset(h->ctc.p, 0, sizeof(float) * (size_t)T * L * h->ldT));
    HIP_TRY(hipMemset(h->nm.p, 0, sizeof(float) * h->ldT));
    HIP_TRY(hipMemset(h->waves.p, 0, sizeof(float) * T * nt * h->ldC));
    HIP_TRY(h->Ucc.alloc((size_t)nC * T * K));
    HIP_TRY(h->xc.alloc(C));
    HIP_TRY(h->yc.alloc(C));
    HIP_TRY(h->iCC.alloc((size_t)nC * C));
    HIP_TRY(h->iU.alloc(T));
    HIP_TRY(h->tw.alloc(lss::kFftN));
    HIP_TRY(h->tw64.alloc(lss::kFftN));
    HIP_TRY(h->hp_dev.alloc(a->hp_len));
    HIP_TRY(hipMemset(h->U.p, 0, sizeof(float) * T * h->ldq));
    HIP_TRY(hipMemcpy2D(h->U.p, sizeof(float) * h->ldq, a->U, sizeof(float) * K * C, sizeof(float) * K * C, T,
                        hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->wpca.p, a->wPCA, sizeof(float) * K * nt, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->xc.p, a->xc, sizeof(float) * C, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->yc.p, a->yc, sizeof(float) * C, hipMemcpyHostToDevice));
    h->h_xc.assign(a->xc, a->xc + C);
    h->h_yc.assign(a->yc, a->yc + C);
    HIP_TRY(hipMemcpy(h->hp_dev.p, a->hp_filter, sizeof(float) * a->hp_len, hipMemcpyHostToDevice));

    // whitening operator (M @ Wrot) in FP64 (kilosort/io.py:787)
    {
        std::vector<float> wm((size_t)C * C);
        if (a->drift) {
            std::vector<double> acc(C);
            for (int i = 0; i < C; ++i) {
                std::fill(acc.begin(), acc.end(), 0.0);
                for (int k = 0; k < C; ++k) {
                    const double mik = a->drift[(size_t)i * C + k];
                    if (mik == 0.0) continue;
                    const float* wr = a->Wrot + (size_t)k * C;
                    for (int j = 0; j < C; ++j) acc[j] += mik * (double)wr[j];
                }
                for (int j = 0; j < C; ++j) wm[(size_t)i * C + j] = (float)acc[j];
            }
        } else {
            std::copy(a->Wrot, a->Wrot + (size_t)C * C, wm.begin());
        }
        HIP_TRY(hipMemset(h->wmat.p, 0, sizeof(float) * C * h->ldm));
        HIP_TRY(hipMemcpy2D(h->wmat.p, sizeof(float) * h->ldm, wm.data(), sizeof(float) * C, sizeof(float) * C, C,
                            hipMemcpyHostToDevice));
        std::vector<double> w64(a->Wrot, a->Wrot + (size_t)C * C);
        HIP_TRY(h->wrot64.alloc((size_t)C * C));
        HIP_TRY(hipMemcpy(h->wrot64.p, w64.data(), sizeof(double) * C * C, hipMemcpyHostToDevice));
    }
    // twiddles exp(-2 pi i m / N)
    {
        std::vector<float2> tw(lss::kFftN);
        const double two_pi = 6.283185307179586476925286766559;
        for (int m = 0; m < lss::kFftN; ++m) {
            double ang = -two_pi * m / lss::kFftN;
            tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
        HIP_TRY(hipMemcpy(h->tw.p, tw.data(), sizeof(float2) * lss::kFftN, hipMemcpyHostToDevice));
        std::vector<double2> tw64(lss::kFftN);
        for (int m = 0; m < lss::kFftN; ++m) {
            double ang = -two_pi * m / lss::kFftN;
            tw64[m] = make_double2(std::cos(ang), std::sin(ang));
        }
        HIP_TRY(hipMemcpy(h->tw64.p, tw64.data(), sizeof(double2) * lss::kFftN, hipMemcpyHostToDevice));
    }
    // iCC / iU / Ucc (prepare_extract, template_matching.py:20-26): the caller's tables where given
    // (the exporter's files, OnlineSpikesV2.cu:561-563), else built from the sites and U
    {
        std::vector<int> iCC;
        if (a->iCC) {
            iCC.assign(a->iCC, a->iCC + (size_t)nC * C);
            for (int v : iCC)
                if (v < 0 || v >= C) return fail(LSS_ERR_ARG, "iCC entry outside [0, n_chan)");
        } else {
            nearest_channels(h->h_xc, h->h_yc, nC, iCC);
        }
        HIP_TRY(hipMemcpy(h->iCC.p, iCC.data(), sizeof(int) * nC * C, hipMemcpyHostToDevice));
        if (lss_status st = refresh_extract(*h)) return st;
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (a->iU) {
            for (int t = 0; t < T; ++t)
                if (a->iU[t] < 0 || a->iU[t] >= C) return fail(LSS_ERR_ARG, "iU entry outside [0, n_chan)");
            HIP_TRY(hipMemcpy(h->iU.p, a->iU, sizeof(int) * T, hipMemcpyHostToDevice));
        }
        if (a->Ucc) {
            HIP_TRY(hipMemcpy(h->Ucc.p, a->Ucc, sizeof(float) * nC * T * K, hipMemcpyHostToDevice));
        } else if (a->iU) {
            // Ucc[j][t][k] = U[t][k][iCC[j][iU[t]]] for the given iU
            std::vector<float> ucc((size_t)nC * T * K);
            for (int j = 0; j < nC; ++j)
                for (int t = 0; t < T; ++t)
                    for (int k = 0; k < K; ++k)
                        ucc[((size_t)j * T + t) * K + k] =
                            a->U[((size_t)t * K + k) * C + iCC[(size_t)j * C + a->iU[t]]];
            HIP_TRY(hipMemcpy(h->Ucc.p, ucc.data(), sizeof(float) * ucc.size(), hipMemcpyHostToDevice));
        }
        h->load.icc_given = a->iCC != nullptr;
        h->load.iu_given = a->iU != nullptr;
        h->load.ucc_given = a->Ucc != nullptr;
    }
    HIP_TRY(h->tables_ws.alloc(lss::tables_workspace_doubles(T, K, C, nt)));
    HIP_TRY(lss::launch_tables(h->U.p, h->ldq, h->wpca.p, T, K, C, nt, h->nm.p, h->waves.p, h->ldC, h->ctc.p, h->ldT,
                               h->tables_ws.p, h->stream));

    // ---- window buffers ----
    HIP_TRY(h->win.alloc((size_t)Wm * C));
    HIP_TRY(h->sums.alloc(C));
    HIP_TRY(h->sum_part.alloc((size_t)lss::channel_sum_blocks(Wm) * C));
    HIP_TRY(h->xcw.alloc((size_t)C * h->ldw));
    HIP_TRY(h->xhp.alloc((size_t)C * h->ldw));
    HIP_TRY(h->xw.alloc((size_t)Wm * h->ldC));
    HIP_TRY(h->xp.alloc((size_t)Wm * h->ldq));
    HIP_TRY(h->B.alloc((size_t)Wm * h->ldT));
    HIP_TRY(hipMemset(h->B.p, 0, sizeof(float) * Wm * h->ldT));
    HIP_TRY(hipMemset(h->xw.p, 0, sizeof(float) * Wm * h->ldC));
    HIP_TRY(h->cfmax.alloc(Wm));
    HIP_TRY(h->imax.alloc(Wm));
    HIP_TRY(h->flags.alloc((size_t)(Wm + 65535) / 65536 * 65536));
    HIP_TRY(hipMemset(h->flags.p, 0, (size_t)(Wm + 65535) / 65536 * 65536));
    HIP_TRY(h->sp_col.alloc(p->max_spikes));
    HIP_TRY(h->sp_tmp.alloc(p->max_spikes));
    HIP_TRY(h->sp_amp.alloc(p->max_spikes));
    HIP_TRY(h->counters.alloc(lss::kCounters));
    HIP_TRY(hipMemset(h->counters.p, 0, sizeof(int) * lss::kCounters));
    HIP_TRY(h->pass_off.alloc(p->max_passes + 2));
    HIP_TRY(h->prow_cf.alloc(2 * (size_t)lss::kPersistMaxW));
    HIP_TRY(h->prow_im.alloc(2 * (size_t)lss::kPersistMaxW));
    HIP_TRY(h->prow_amp.alloc(2 * (size_t)lss::kPersistMaxW));
    HIP_TRY(h->pbar.alloc((size_t)lss::kPersistMaxGrid * lss::kBarStride));
    HIP_TRY(hipMemset(h->pbar.p, 0, sizeof(int) * lss::kPersistMaxGrid * lss::kBarStride));
    HIP_TRY(h->tF.alloc((size_t)p->max_spikes * nC * K));
    HIP_TRY(h->xs.alloc(p->max_spikes));
    HIP_TRY(h->ys.alloc(p->max_spikes));
    HIP_TRY(hipHostMalloc((void**)&h->h_counters, sizeof(int) * 8, hipHostMallocDefault));
    for (auto& e : h->ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->alive = true;
    *out = h.release();
    return LSS_OK;
}

Handle* as_handle(void* p) {
    Handle* h = static_cast<Handle*>(p);
    if (h && h->pending) {
        (void)hipEventSynchronize(h->pending);
        h->pending = nullptr;
        h->touched = true;
    }
    return h;
}

}  // namespace

namespace lss {
lss_status set_error(lss_status st, const char* msg) { return fail(st, msg); }
}  // namespace lss

namespace {

lss_status check_window(Handle& h, int64_t W) {
    if (W < 2 * (int64_t)h.nt + 2) return fail(LSS_ERR_ARG, "window shorter than 2*nt+2");
    if (W > h.Wmax) return fail(LSS_ERR_ARG, "window longer than max_window");
    return LSS_OK;
}

lss_status run_car(Handle& h, const int16_t* d_win, int64_t W) {
    HIP_TRY(lss::launch_channel_sums(d_win, W, h.C, h.sum_part.p, h.sums.p, h.stream));
    HIP_TRY(lss::launch_car_transpose(d_win, W, h.C, h.sums.p, h.prm.do_car, h.xcw.p, h.ldw, h.stream));
    return LSS_OK;
}

lss_status run_highpass(Handle& h, const float* in, float* out, int64_t W) {
    const Handle::Spec* sp = nullptr;
    if (lss_status st = get_spectrum(h, W, false, &sp)) return st;
    HIP_TRY(lss::launch_highpass(in, h.ldw, out, h.ldw, h.C, W, sp->lo, sp->hi, sp->p, h.tw.p, h.stream));
    return LSS_OK;
}

lss_status run_whiten(Handle& h, const float* in_cw, float* out_wc, int64_t W) {
    // Xw[w][c'] = sum_c X[c][w] Wm[c'][c]
    if (h.prm.precision == LSS_PREC_FP32)
        HIP_TRY(lss::launch_gemm_f32(in_cw, h.ldw, 1, h.wmat.p, h.ldm, out_wc, h.ldC, (int)W, h.C, h.C, h.stream,
                                     h.tile_whiten));
    else
        HIP_TRY(lss::launch_gemm_lp(h.prm.precision, in_cw, h.ldw, 1, h.wmat.p, h.ldm, out_wc, h.ldC, (int)W, h.C, h.C,
                                    h.stream, h.tile_whiten));
    return LSS_OK;
}

lss_status run_scores(Handle& h, const float* x_wc, float* b_wt, int64_t W, hipEvent_t mid = nullptr) {
    // the split-precision paths project on the matrix cores with the same split as their score GEMM
    if (h.prm.precision == LSS_PREC_BF16X3 || h.prm.precision == LSS_PREC_FP16X3)
        HIP_TRY(lss::launch_project_x3(h.prm.precision, x_wc, h.ldC, W, h.C, h.wpca.p, h.K, h.nt, h.xp.p, h.ldq,
                                       h.stream, h.proj_tile));
    else
        HIP_TRY(lss::launch_project(x_wc, h.ldC, W, h.C, h.wpca.p, h.K, h.nt, h.xp.p, h.ldq, h.stream, h.proj_tile));
    if (mid) HIP_TRY(hipEventRecord(mid, h.stream));
    const int KC = h.K * h.C;
    if (h.prm.precision == LSS_PREC_FP32)
        HIP_TRY(lss::launch_gemm_f32(h.xp.p, h.ldq, 0, h.U.p, h.ldq, b_wt, h.ldT, (int)W, h.T, KC, h.stream,
                                     h.tile_scores));
    else
        HIP_TRY(lss::launch_gemm_lp(h.prm.precision, h.xp.p, h.ldq, 0, h.U.p, h.ldq, b_wt, h.ldT, (int)W, h.T, KC,
                                    h.stream, h.tile_scores));
    return LSS_OK;
}

// workgroups of the PC-feature kernel (four spikes each per step, grid-stride): 2048 gives most
// spikes a wavefront of their own (36 us for ~11 k spikes; 1024: 42 us, 4096: 37 us)
constexpr int kFeatureGrid = 2048;

lss::PursuitArgs pursuit_args(Handle& h, float* B, float* X, int64_t W) {
    h.last_W = W;
    lss::PursuitArgs a{};
    a.B = B; a.X = X; a.nm = h.nm.p; a.ctc = h.ctc.p; a.waves = h.waves.p;
    a.pass_off = h.pass_off.p;
    a.prow_cf = h.prow_cf.p; a.prow_im = h.prow_im.p; a.prow_amp = h.prow_amp.p; a.bar = h.pbar.p;
    a.cfmax = h.cfmax.p; a.imax = h.imax.p; a.flags = h.flags.p;
    a.sp_col = h.sp_col.p; a.sp_tmp = h.sp_tmp.p; a.sp_amp = h.sp_amp.p; a.counters = h.counters.p;
    a.W = W; a.ldT = h.ldT; a.ldC = h.ldC; a.T = h.T; a.C = h.C; a.K = h.K; a.nt = h.nt; a.cap = h.prm.max_spikes;
    // Th**2 in double, rounded to FP32 as the reference's float32 comparison rounds it
    a.th2 = (float)(h.prm.th_learned * h.prm.th_learned);
    return a;
}

// Pursuit passes [p0, p1) on the handle's stream with no host round trip (counts stay on the
// device); pass 0 also resets the counters and scores every column.
lss_status enqueue_passes(Handle& h, const lss::PursuitArgs& a, int p0, int p1) {
    if (p0 == 0) {
        HIP_TRY(lss::launch_reset_counters(h.counters.p, h.stream, h.pbar.p));
        HIP_TRY(lss::launch_score_all(a, h.stream));
    }
    for (int p = p0; p < p1; ++p) {
        HIP_TRY(lss::launch_peaks(a, h.stream));
        if (h.prof_update) {
            lss::PursuitArgs b = a;
            // stamps only while the buffer matches the grid (lss_profile_update sizes it)
            if (h.upd_ts.n == (size_t)h.prm.max_passes * 2 * h.upd_grid) b.ts = h.upd_ts.p + (size_t)p * 2 * h.upd_grid;
            HIP_TRY(lss::launch_update(b, h.upd_grid, h.stream, h.upd_ev[2 * p], h.upd_ev[2 * p + 1], h.upd_cpw));
        } else
            HIP_TRY(lss::launch_update(a, h.upd_grid, h.stream, nullptr, nullptr, h.upd_cpw));
    }
    return LSS_OK;
}

bool use_persistent(const Handle& h, int64_t W) {
    return h.pers_grid > 0 && W <= lss::kPersistMaxW && lss::persistent_supports(h.T);
}

// The pursuit of one window: short windows run their first pers_first passes (the wide ones) as
// per-pass kernels and the rest in the persistent kernel (grid capped at max_grid when > 0);
// others, and short windows whose score rows do not fit the LDS of that grid, get passes [0, p1)
// as per-pass kernels.
lss_status enqueue_pursuit(Handle& h, const lss::PursuitArgs& a, int p1, int max_grid) {
    int g = 0;
    if (use_persistent(h, a.W))
        HIP_TRY(lss::persistent_grid(a, max_grid > 0 ? std::min(h.pers_grid, max_grid) : h.pers_grid, &g));
    if (g > 0) {
        const int p0 = std::min(h.pers_first, h.prm.max_passes - 1);
        if (p0 > 0) {
            if (lss_status st = enqueue_passes(h, a, 0, p0)) return st;
        } else {
            HIP_TRY(lss::launch_reset_counters(h.counters.p, h.stream, h.pbar.p));
        }
        lss::PursuitArgs b = a;
        if (h.prof_update && h.pers_ts.p) b.ts = h.pers_ts.p;
        HIP_TRY(lss::launch_pursuit_persistent(b, g, p0, h.prm.max_passes, h.stream, &h.pers_used));
        h.pers_p0 = p0;
        return LSS_OK;
    }
    h.pers_p0 = -1;
    return enqueue_passes(h, a, 0, p1);
}

// All max_passes passes, then the residual of every pass.
lss_status run_pursuit(Handle& h, float* B, float* X, int64_t W) {
    const lss::PursuitArgs a = pursuit_args(h, B, X, W);
    if (lss_status st = enqueue_pursuit(h, a, h.prm.max_passes, 0)) return st;
    HIP_TRY(lss::launch_residual(a, h.stream));
    return LSS_OK;
}

lss::FeatureArgs feature_args(const Handle& h, const float* X, int64_t W) {
    lss::FeatureArgs f{};
    f.X = X; f.wpca = h.wpca.p; f.iCC = h.iCC.p; f.iU = h.iU.p; f.Ucc = h.Ucc.p;
    f.xc = h.xc.p; f.yc = h.yc.p; f.sp_col = h.sp_col.p; f.sp_tmp = h.sp_tmp.p; f.sp_amp = h.sp_amp.p;
    f.counters = h.counters.p;
    f.tF = h.tF.p; f.xs = h.xs.p; f.ys = h.ys.p;
    f.n = 0; f.cap = h.prm.max_spikes; f.C = h.C; f.T = h.T; f.K = h.K; f.nt = h.nt; f.nC = h.nC;
    f.W = W; f.ldC = h.ldC;
    return f;
}

lss_status run_features(Handle& h, const float* X, int64_t W) {
    HIP_TRY(lss::launch_features(feature_args(h, X, W), kFeatureGrid, h.stream));
    return LSS_OK;
}

// Device part of one window up to the scores (no events, no host synchronisation: capturable).
// Preprocessing and scores of the window d_win into xw (whitened data) and B (scores).
lss_status enqueue_pre(Handle& h, const Handle::Spec& sp, const int16_t* d_win, int64_t W, float* xw, float* B) {
    if (lss_status st = run_car(h, d_win, W)) return st;
    HIP_TRY(lss::launch_highpass(h.xcw.p, h.ldw, h.xhp.p, h.ldw, h.C, W, sp.lo, sp.hi, sp.p, h.tw.p, h.stream));
    if (lss_status st = run_whiten(h, h.xhp.p, xw, W)) return st;
    return run_scores(h, xw, B, W);
}

// Pinned, device-mapped, coherent host buffers that the staging kernel writes directly.  Windows
// with more spikes than this are read back by copies (emit_spikes).
constexpr int kStageCap = 1 << 16;

lss_status ensure_stage(Handle& h) {
    if (h.stage_cap) return LSS_OK;
    const int cap = std::min(h.prm.max_spikes, kStageCap);
    const int FK = h.nC * h.K;
    const unsigned fl = hipHostMallocMapped | hipHostMallocCoherent;
    // buffers left by an earlier attempt that failed part way are kept (the destructor frees them)
    auto pin = [&](void** p, size_t bytes) { return *p ? hipSuccess : hipHostMalloc(p, bytes, fl); };
    HIP_TRY(pin((void**)&h.hs_counters, sizeof(int) * 8));
    HIP_TRY(pin((void**)&h.hs_col, sizeof(int) * cap));
    HIP_TRY(pin((void**)&h.hs_tmp, sizeof(int) * cap));
    HIP_TRY(pin((void**)&h.hs_amp, sizeof(float) * cap));
    HIP_TRY(pin((void**)&h.hs_tF, sizeof(float) * (size_t)cap * FK));
    HIP_TRY(pin((void**)&h.hs_xs, sizeof(float) * cap));
    HIP_TRY(pin((void**)&h.hs_ys, sizeof(float) * cap));
    lss::StageArgs& g = h.stage;
    g.counters = h.counters.p;
    g.col = h.sp_col.p; g.tmp = h.sp_tmp.p; g.amp = h.sp_amp.p; g.tF = h.tF.p; g.xs = h.xs.p; g.ys = h.ys.p;
    HIP_TRY(hipHostGetDevicePointer((void**)&g.h_counters, h.hs_counters, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&g.h_col, h.hs_col, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&g.h_tmp, h.hs_tmp, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&g.h_amp, h.hs_amp, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&g.h_tF, h.hs_tF, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&g.h_xs, h.hs_xs, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&g.h_ys, h.hs_ys, 0));
    g.cap = cap;
    g.FK = FK;
    h.stage_cap = cap;
    return LSS_OK;
}

// Counters and the first stage_cap spikes to pinned host memory by one kernel that moves only the
// spikes found (capturable).
lss_status enqueue_readback(Handle& h) {
    HIP_TRY(lss::launch_stage_spikes(h.stage, h.stream));
    return LSS_OK;
}

constexpr int kPersStamps = 2048;   // persistent-pursuit workgroups measurement mode can stamp
const char* const kBarrierTimeout =
    "persistent pursuit: grid barrier timed out (workgroups not co-resident); results are invalid";

// Spike count and pass count of the last pursuit (synchronises the stream).
// The error flags and counts of a finished window, from a host copy c of the device counters.
lss_status check_counts(Handle& h, const int* c, int* n, int* passes) {
    if (c[7]) return fail(LSS_ERR_STATE, kBarrierTimeout);
    if (c[2]) return fail(LSS_ERR_CAPACITY, "spike capacity (max_spikes) exceeded");
    if (c[5])
        return fail(LSS_ERR_CAPACITY, "more than 1024 spikes of a single pass (tied peaks) reach one 16-column "
                                      "block of the residual; the residual is incomplete");
    *n = c[0] + c[1];
    *passes = c[4];
    if (h.prof_update) {
        // per-pass update launches of this window: all max_passes, or the first pers_p0 when the
        // persistent pursuit took the rest (the other events were not recorded)
        const int ran = h.pers_p0 >= 0 ? std::min(h.pers_p0, h.prm.max_passes) : h.prm.max_passes;
        h.upd_ms = 0.f;
        h.upd_launches = h.prm.max_passes;
        h.upd_each.assign(h.prm.max_passes, 0.f);
        h.upd_spikes.assign(h.prm.max_passes, 0);
        for (int p = 0; p < ran; ++p) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, h.upd_ev[2 * p], h.upd_ev[2 * p + 1]));
            h.upd_ms += ms;
            h.upd_each[p] = ms;
        }
        // GPU-clock duration of launch p: last workgroup end - first workgroup start (10 ns ticks)
        const size_t nts = (size_t)h.prm.max_passes * 2 * h.upd_grid;
        h.upd_gpu.assign(h.prm.max_passes, 0.f);
        if (h.upd_ts.n == nts) {
            std::vector<unsigned long long> span(2 * (size_t)h.prm.max_passes);
            HIP_TRY(lss::launch_stamp_span(h.upd_ts.p, h.prm.max_passes, h.upd_grid, h.upd_span.p, h.stream));
            HIP_TRY(hipMemcpyAsync(span.data(), h.upd_span.p, sizeof(unsigned long long) * span.size(),
                                   hipMemcpyDeviceToHost, h.stream));
            HIP_TRY(hipStreamSynchronize(h.stream));
            for (int p = 0; p < ran; ++p) {
                const unsigned long long t0 = span[2 * p], t1 = span[2 * p + 1];
                h.upd_gpu[p] = t1 > t0 ? (float)((double)(t1 - t0) * 1e-5) : 0.f;
            }
        }
        // launch p applied the spikes of pass p (passes that found spikes come first)
        std::vector<int> off(h.prm.max_passes + 2, 0);
        HIP_TRY(hipMemcpyAsync(off.data(), h.pass_off.p, sizeof(int) * (h.prm.max_passes + 2), hipMemcpyDeviceToHost,
                               h.stream));
        HIP_TRY(hipStreamSynchronize(h.stream));
        // per-pass launches ran passes [0, pers_p0) when the persistent pursuit took the rest
        const int pp = h.pers_p0 >= 0 ? std::min(h.pers_p0, *passes) : *passes;
        for (int p = 0; p < pp && p < h.prm.max_passes; ++p) h.upd_spikes[p] = off[p + 1] - off[p];
        h.pers_ms = 0.f;
        h.pers_spikes = h.pers_passes = 0;
        if (h.pers_p0 >= 0 && h.pers_used > 0 && h.pers_used <= kPersStamps) {
            unsigned long long span[2] = {0, 0};
            HIP_TRY(lss::launch_stamp_span(h.pers_ts.p, 1, h.pers_used, h.pers_span.p, h.stream));
            HIP_TRY(hipMemcpyAsync(span, h.pers_span.p, sizeof(span), hipMemcpyDeviceToHost, h.stream));
            HIP_TRY(hipStreamSynchronize(h.stream));
            h.pers_ms = span[1] > span[0] ? (float)((double)(span[1] - span[0]) * 1e-5) : 0.f;
            h.pers_spikes = (*passes > h.pers_p0) ? *n - off[h.pers_p0] : 0;
            h.pers_passes = std::max(0, *passes - h.pers_p0);
        }
    }
    return LSS_OK;
}

lss_status read_counts(Handle& h, int* n, int* passes) {
    HIP_TRY(hipMemcpyAsync(h.h_counters, h.counters.p, sizeof(int) * 8, hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    return check_counts(h, h.h_counters, n, passes);
}

// Host-side emission of n device-order spikes from host copies of the spike arrays.
lss_status emit_from(Handle& h, int n, int passes, int64_t window_start, lss_spikes* out, const int* col,
                     const int* tmp, const float* amp, const float* tF, const float* xs, const float* ys) {
    const int FK = h.nC * h.K;
    int64_t k = 0;
    const int64_t shift = window_start - h.nt / 2 + h.prm.nt0min;   // template_matching.py:68
    for (int i = 0; i < n; ++i) {
        const int64_t smp = shift + col[i];
        if (h.prm.drop_negative && smp < 0) continue;   // template_matching.py:53-59
        if (k >= out->capacity) return fail(LSS_ERR_CAPACITY, "output capacity exceeded");
        if (out->sample) out->sample[k] = smp;
        if (out->template_id) out->template_id[k] = tmp[i];
        if (out->amplitude) out->amplitude[k] = amp[i];
        if (out->column) out->column[k] = col[i];
        if (out->features) std::memcpy(out->features + k * FK, tF + (size_t)i * FK, sizeof(float) * FK);
        if (out->x) out->x[k] = xs[i];
        if (out->y) out->y[k] = ys[i];
        ++k;
    }
    out->n = k;
    out->passes = passes;
    return LSS_OK;
}

lss_status emit_spikes(Handle& h, int n, int passes, int64_t window_start, lss_spikes* out) {
    out->n = 0;
    out->passes = passes;
    if (n == 0) return LSS_OK;
    const int FK = h.nC * h.K;
    std::vector<int> col(n), tmp(n);
    std::vector<float> amp(n), tF((size_t)n * FK), xs(n), ys(n);
    HIP_TRY(hipMemcpyAsync(col.data(), h.sp_col.p, sizeof(int) * n, hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipMemcpyAsync(tmp.data(), h.sp_tmp.p, sizeof(int) * n, hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipMemcpyAsync(amp.data(), h.sp_amp.p, sizeof(float) * n, hipMemcpyDeviceToHost, h.stream));
    if (out->features) HIP_TRY(hipMemcpyAsync(tF.data(), h.tF.p, sizeof(float) * n * FK, hipMemcpyDeviceToHost, h.stream));
    if (out->x) HIP_TRY(hipMemcpyAsync(xs.data(), h.xs.p, sizeof(float) * n, hipMemcpyDeviceToHost, h.stream));
    if (out->y) HIP_TRY(hipMemcpyAsync(ys.data(), h.ys.p, sizeof(float) * n, hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    return emit_from(h, n, passes, window_start, out, col.data(), tmp.data(), amp.data(), tF.data(), xs.data(),
                     ys.data());
}

// ------------------------------------------------------------------------------------------------
// Multi-probe streaming batcher.  Probes share one sample clock; every step sorts the next `batch`
// samples of every probe.  A probe's window is samples [t0 - nt, t0 + batch + nt) (Kilosort
// padding: samples before 0 repeat sample 0), so consecutive windows report disjoint sample
// ranges.  Device work per step: segment A = window upload, preprocessing, scores and the first P1
// pursuit passes of every probe as parallel graph branches; one host check of the converged flags;
// extra passes (non-graph) for probes that need them; segment C = residual, features and the
// readback of every probe.  With use_graph, A and C are captured once and replayed.  When every
// probe runs the persistent pursuit (all passes in A), C joins A and a step is one graph.
//
// Running ahead (one-graph steps only): when the next batch's samples are already here and no
// refinement follows the step, the step also preprocesses the next batch (window, whitening,
// scores) into the handles' second window sets: segment N, a second graph on its own stream (one
// branch per probe) beside segment A, so that the next step starts with the pursuit.  Each step still reports only
// its own batch, and the results are the same as without running ahead (the preprocessing of a
// batch reads only its samples and the templates, which do not change between the two steps).  A
// step that ran ahead is discarded for a probe whose preprocessing inputs changed in between
// (lss_set_drift, lss_refine_templates: Handle::pre_version); the batch is then preprocessed again.
// ------------------------------------------------------------------------------------------------
struct Probe {
    Handle* h = nullptr;
    int mode = 0;                 // 0 nothing pushed yet, 1 host samples, 2 device samples
    std::vector<int16_t> buf;     // host mode: received samples [base, base + rows) as [row][C]
    std::deque<std::pair<const int16_t*, int64_t>> dchunks;   // device mode: (pointer, first sample)
    std::deque<int64_t> dlen;     // rows of each device chunk
    int64_t base = 0;
    int64_t recv = 0;
    int16_t* pin = nullptr;       // pinned window [W][C]
    int16_t* pin_alt = nullptr;   // the next batch's (host mode, running ahead)
    int64_t W = 0;
    int P1 = 0;
    int max_grid = 0;             // persistent-pursuit workgroups, so that all probes' fit at once
    const Handle::Spec* spec = nullptr;   // the window's highpass spectrum (pinned in the handle)
    hipEvent_t done = nullptr;
    // running ahead: the next batch's preprocessing stream and its end, the handle's pre_version
    // when it ran
    hipStream_t s2 = nullptr;
    hipEvent_t done2 = nullptr;
    uint64_t ahead_version = 0;
    // refinement in the background: the job, the step count at whose step start it is adopted
    // (-1: none pending), whether it has been joined, its status, message, result and info; the
    // event after the copy of its inputs
    std::thread rjob;
    int64_t rjob_adopt = -1;
    bool rjob_joined = false;
    lss_status rjob_st = LSS_OK;
    std::string rjob_err;
    bool rjob_adopted = false;
    std::array<double, 4> rjob_info{};
    hipEvent_t rsnap = nullptr;
};

// Preprocessing lanes of a grouped step: with the pursuit tail's stream, no more hardware queues
// than the runtime gives a process by default (4).
constexpr int kPreLanes = 3;

struct Batcher {
    std::vector<Probe> pr;
    int batch = 0;
    bool use_graph = false;
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr;
    // running ahead: the next batch's preprocessing (segment N) on its own stream, as a second graph
    // beside segment A; by swapped
    hipStream_t main2 = nullptr;
    hipEvent_t fork2 = nullptr, joinA = nullptr;
    hipEvent_t staged = nullptr;   // b.main's work before segment A: adopted templates, gathered windows
    hipGraph_t graphsN[2] = {};
    hipGraphExec_t execsN[2] = {};
    hipGraphExec_t gA = nullptr, gC = nullptr;
    hipGraph_t graphA = nullptr, graphC = nullptr;
    int64_t t0 = 0;
    int64_t batches = 0;
    float last_ms = 0.f;
    bool single = false;     // every probe runs all its passes in segment A: one graph per step
    bool run_ahead = true;   // preprocess the next batch during a step when its samples are here
    bool ahead = false;      // the last step preprocessed this step's batch into the second sets
    int64_t ran_ahead = 0;   // steps that preprocessed their next batch
    int swapped = 0;         // the window sets have been swapped an odd number of times
    bool cur_pre = true, next_pre = false;   // the step preprocesses this batch / the next batch
    // one-graph steps, by (cur_pre, swapped): cur_pre | swapped << 1
    hipGraph_t graphs[4] = {};
    hipGraphExec_t execs[4] = {};
    int refine_every = 0;    // template refinement round every this many batches (0 = never)
    int refine_next = -1;    // probe the current refinement round refines next (-1: no round)
    std::vector<std::array<double, 4>> refine_info;   // per probe, last refinement (lss_refine_templates)
    float part_ms[6] = {};   // assemble, segment A, extra passes, segment C, emit, total
    // grouped pursuit tail: every probe's counter reset, persistent pursuit, residual, features and
    // readback in one launch each on the main stream (lss_stream_set_grouped); group: -1 not yet
    // decided, 0 per-probe branches, 1 grouped.  Member descriptors by swapped (device and host),
    // each probe's persistent grid and max_passes.
    bool grouped = true;
    int group = -1;
    DevBuf<lss::GroupMember> gdev[2];
    std::vector<lss::GroupMember> ghost[2];
    int ggrid[lss::kGroupMax] = {}, gmp[lss::kGroupMax] = {};
    // grouped steps: the preprocessing lanes (probes of each, in order), each on a stream with a
    // hardware queue of its own, with its end events (this batch, the next batch); graphs of the
    // lanes by (next batch << 1 | swapped) and of the pursuit tail by swapped
    int nlanes = 0;
    std::vector<int> lanes[kPreLanes];
    hipStream_t lane_s[kPreLanes] = {};
    hipEvent_t lane_ev[kPreLanes] = {}, lane_ev2[kPreLanes] = {};
    hipGraph_t gpre[4][kPreLanes] = {};
    hipGraphExec_t xpre[4][kPreLanes] = {};
    hipGraph_t gtail[2] = {};
    hipGraphExec_t xtail[2] = {};
    // measurement mode of grouped steps (lss_stream_profile): the tail launched directly with
    // descriptors whose members stamp their workgroups' GPU-clock start and end into one buffer
    // (member i from workgroup first[i] on); the last step's span of the grouped persistent launch
    bool profile = false;
    // pipelined steps (lss_stream_set_pipelined): a step whose next batch it has preprocessed also
    // launches that batch's pursuit tail before it returns (early); the next step uses it unless a
    // handle was touched or its preprocessing inputs changed in between (tail_inflight: it does)
    bool pipelined = true;
    bool early = false, tail_inflight = false;
    hipEvent_t early_ev = nullptr;
    int64_t pipelined_steps = 0;
    DevBuf<lss::GroupMember> gprof[2];
    DevBuf<unsigned long long> pts, pspan;
    double prof_us = 0.0;
    ~Batcher() {
        for (auto& p : pr)
            if (p.h && early_ev && p.h->pending == early_ev) p.h->pending = nullptr;
        if (early_ev) (void)hipEventDestroy(early_ev);
        for (int k = 0; k < 4; ++k)
            for (int l = 0; l < kPreLanes; ++l) {
                if (xpre[k][l]) (void)hipGraphExecDestroy(xpre[k][l]);
                if (gpre[k][l]) (void)hipGraphDestroy(gpre[k][l]);
            }
        for (int k = 0; k < 2; ++k) {
            if (xtail[k]) (void)hipGraphExecDestroy(xtail[k]);
            if (gtail[k]) (void)hipGraphDestroy(gtail[k]);
        }
        for (int l = 0; l < kPreLanes; ++l) {
            if (lane_ev[l]) (void)hipEventDestroy(lane_ev[l]);
            if (lane_ev2[l]) (void)hipEventDestroy(lane_ev2[l]);
            if (lane_s[l]) (void)hipStreamDestroy(lane_s[l]);
        }
        if (gA) (void)hipGraphExecDestroy(gA);
        if (gC) (void)hipGraphExecDestroy(gC);
        if (graphA) (void)hipGraphDestroy(graphA);
        if (graphC) (void)hipGraphDestroy(graphC);
        for (int i = 0; i < 4; ++i) {
            if (execs[i]) (void)hipGraphExecDestroy(execs[i]);
            if (graphs[i]) (void)hipGraphDestroy(graphs[i]);
        }
        for (int i = 0; i < 2; ++i) {
            if (execsN[i]) (void)hipGraphExecDestroy(execsN[i]);
            if (graphsN[i]) (void)hipGraphDestroy(graphsN[i]);
        }
        if (fork2) (void)hipEventDestroy(fork2);
        if (joinA) (void)hipEventDestroy(joinA);
        if (staged) (void)hipEventDestroy(staged);
        if (main2) (void)hipStreamDestroy(main2);
        for (auto& p : pr) {
            if (p.pin) (void)hipHostFree(p.pin);
            if (p.pin_alt) (void)hipHostFree(p.pin_alt);
            if (p.done) (void)hipEventDestroy(p.done);
            if (p.rjob.joinable()) {
                p.rjob.join();
                p.h->rf_busy = false;
            }
            if (p.rsnap) (void)hipEventDestroy(p.rsnap);
            if (p.done2) (void)hipEventDestroy(p.done2);
            if (p.s2) (void)hipStreamDestroy(p.s2);
            if (p.spec) --const_cast<Handle::Spec*>(p.spec)->pins;
        }
        if (fork) (void)hipEventDestroy(fork);
        if (main) (void)hipStreamDestroy(main);
    }
};

// Upload (host mode) and preprocessing of one of probe p's window sets.
lss_status probe_pre(Probe& p, bool alt) {
    Handle& h = *p.h;
    int16_t* win = alt ? h.win_alt.p : h.win.p;
    if (p.mode == 1)
        HIP_TRY(hipMemcpyAsync(win, alt ? p.pin_alt : p.pin, sizeof(int16_t) * p.W * h.C, hipMemcpyHostToDevice,
                               h.stream));
    return enqueue_pre(h, *p.spec, win, p.W, alt ? h.xw_alt.p : h.xw.p, alt ? h.B_alt.p : h.B.p);
}

// Probe part of segment A: upload, preprocessing (unless the last step ran it ahead), scores and
// the first P1 passes (all passes for a persistent pursuit); the counters go to the host only when
// extra passes may follow.
lss_status probe_A(Probe& p, bool pre = true) {
    Handle& h = *p.h;
    if (pre)
        if (lss_status st = probe_pre(p, false)) return st;
    const lss::PursuitArgs a = pursuit_args(h, h.B.p, h.xw.p, p.W);
    if (lss_status st = enqueue_pursuit(h, a, p.P1, p.max_grid)) return st;
    if (p.P1 < h.prm.max_passes)
        HIP_TRY(hipMemcpyAsync(h.hs_counters, h.counters.p, sizeof(int) * 8, hipMemcpyDeviceToHost, h.stream));
    return LSS_OK;
}

// Probe part of segment C: residual, features, readback.
lss_status probe_C(Probe& p) {
    Handle& h = *p.h;
    const lss::PursuitArgs a = pursuit_args(h, h.B.p, h.xw.p, p.W);
    HIP_TRY(lss::launch_residual(a, h.stream));
    if (lss_status st = run_features(h, h.xw.p, p.W)) return st;
    return enqueue_readback(h);
}

// The pursuit tail of a grouped step on the main stream, one launch per stage for all probes:
// counter resets, persistent pursuits, residuals, features with the readbacks.
lss_status segment_tail(Batcher& b) {
    const int n = (int)b.pr.size();
    const lss::GroupMember* d = b.profile ? b.gprof[b.swapped].p : b.gdev[b.swapped].p;
    const lss::GroupMember* m = b.ghost[b.swapped].data();
    HIP_TRY(lss::launch_group_reset(d, n, b.main));
    HIP_TRY(lss::launch_group_persistent(d, m, b.ggrid, b.gmp, n, b.main));
    HIP_TRY(lss::launch_group_residual(d, m, n, b.main));
    HIP_TRY(lss::launch_group_features(d, m, n, b.main));   // also writes the readback staging
    for (int i = 0; i < n; ++i) {
        b.pr[i].h->pers_p0 = 0;
        b.pr[i].h->pers_used = b.ggrid[i];
    }
    return LSS_OK;
}

// Lane l's preprocessing of its probes' current (alt = false) or second (alt = true) window sets,
// one probe after another on the lane's stream.
lss_status lane_pre(Batcher& b, int l, bool alt) {
    for (int i : b.lanes[l]) {
        Probe& p = b.pr[i];
        Handle& h = *p.h;
        hipStream_t s = b.lane_s[l];
        std::swap(h.stream, s);   // the probe functions enqueue on h.stream
        const lss_status st = probe_pre(p, alt);
        std::swap(h.stream, s);
        if (st != LSS_OK) return st;
    }
    return LSS_OK;
}

// Captures what f enqueues on stream s into a graph and instantiates it.
template <class F>
lss_status capture_on(hipStream_t s, F f, hipGraph_t* g, hipGraphExec_t* ge) {
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    const lss_status st = f();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (st != LSS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return fail(LSS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    *g = graph;
    HIP_TRY(hipGraphInstantiate(ge, graph, nullptr, nullptr, 0));
    return LSS_OK;
}

lss_status run_lane(Batcher& b, int l, bool alt) {
    if (!b.use_graph) return lane_pre(b, l, alt);
    const int key = (alt ? 2 : 0) | b.swapped;
    if (!b.xpre[key][l])
        if (lss_status st = capture_on(b.lane_s[l], [&] { return lane_pre(b, l, alt); }, &b.gpre[key][l],
                                       &b.xpre[key][l]))
            return st;
    HIP_TRY(hipGraphLaunch(b.xpre[key][l], b.lane_s[l]));
    return LSS_OK;
}

// The pursuit tail of a grouped step on the main stream (its graph, or direct launches).
lss_status launch_tail(Batcher& b) {
    if (b.use_graph && !b.profile) {
        if (!b.xtail[b.swapped])
            if (lss_status st = capture_on(b.main, [&] { return segment_tail(b); }, &b.gtail[b.swapped],
                                           &b.xtail[b.swapped]))
                return st;
        HIP_TRY(hipGraphLaunch(b.xtail[b.swapped], b.main));
        return LSS_OK;
    }
    return segment_tail(b);
}

// The device work of a grouped step.  The lanes preprocess this batch when the last step did not
// (the tail waits for them), the tail runs on the main stream, and the lanes preprocess the next
// batch beside it (running ahead).  Every lane starts after the main stream's work so far (the
// gathered windows, templates adopted at the step's start); on a lane the next batch's
// preprocessing follows this batch's, which shares the probes' intermediate buffers.
lss_status grouped_step(Batcher& b) {
    // with this batch's tail already in flight on the main stream the windows were gathered on lane
    // 0 (nothing else on the main stream precedes this step's preprocessing then)
    HIP_TRY(hipEventRecord(b.staged, b.tail_inflight ? b.lane_s[0] : b.main));
    for (int l = 0; l < b.nlanes; ++l)
        if (b.cur_pre || b.next_pre) HIP_TRY(hipStreamWaitEvent(b.lane_s[l], b.staged, 0));
    if (b.cur_pre)
        for (int l = 0; l < b.nlanes; ++l) {
            if (lss_status st = run_lane(b, l, false)) return st;
            HIP_TRY(hipEventRecord(b.lane_ev[l], b.lane_s[l]));
            HIP_TRY(hipStreamWaitEvent(b.main, b.lane_ev[l], 0));
        }
    if (!b.tail_inflight)
        if (lss_status st = launch_tail(b)) return st;
    if (b.next_pre)
        for (int l = 0; l < b.nlanes; ++l) {
            if (lss_status st = run_lane(b, l, true)) return st;
            HIP_TRY(hipEventRecord(b.lane_ev2[l], b.lane_s[l]));
        }
    if (b.next_pre)
        for (int l = 0; l < b.nlanes; ++l) HIP_TRY(hipEventSynchronize(b.lane_ev2[l]));
    if (b.profile && !b.tail_inflight) {
        int total = 0;
        for (size_t i = 0; i < b.pr.size(); ++i) total += b.ggrid[i];
        unsigned long long span[2] = {0, 0};
        HIP_TRY(lss::launch_stamp_span(b.pts.p, 1, total, b.pspan.p, b.main));
        HIP_TRY(hipMemcpyAsync(span, b.pspan.p, sizeof(span), hipMemcpyDeviceToHost, b.main));
        HIP_TRY(hipStreamSynchronize(b.main));
        b.prof_us = span[1] > span[0] ? (double)(span[1] - span[0]) * 1e-2 : 0.0;   // 100 MHz clock
    }
    HIP_TRY(hipStreamSynchronize(b.main));
    return LSS_OK;
}

// A stream with a hardware queue of its own: a stream with a compute-unit mask (here every CU of
// the device) gets a queue that the runtime's shared pool never hands to another stream.  The
// batcher's streams are blocking streams (ordered after earlier work of the legacy default stream,
// which device samples pushed from torch's default stream rely on); should the runtime make this
// one non-blocking, an ordinary blocking stream on the shared pool replaces it.
lss_status own_queue_stream(hipStream_t* s) {
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    std::vector<uint32_t> mask((cus + 31) / 32, 0xFFFFFFFFu);
    if (cus % 32) mask.back() = (1u << (cus % 32)) - 1u;
    HIP_TRY(hipExtStreamCreateWithCUMask(s, (uint32_t)mask.size(), mask.data()));
    unsigned flags = 0;
    HIP_TRY(hipStreamGetFlags(*s, &flags));
    if (flags & hipStreamNonBlocking) {
        HIP_TRY(hipStreamDestroy(*s));
        HIP_TRY(hipStreamCreate(s));
    }
    return LSS_OK;
}

// The lanes of grouped steps: the largest probes (by C*T) one each, the smallest sharing the last.
lss_status make_lanes(Batcher& b) {
    const int n = (int)b.pr.size();
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        return (int64_t)b.pr[x].h->C * b.pr[x].h->T > (int64_t)b.pr[y].h->C * b.pr[y].h->T;
    });
    if (!b.staged) HIP_TRY(hipEventCreateWithFlags(&b.staged, hipEventDisableTiming));
    b.nlanes = std::min(n, kPreLanes);
    for (int r = 0; r < n; ++r) b.lanes[std::min(r, kPreLanes - 1)].push_back(ord[r]);
    for (int l = 0; l < b.nlanes; ++l) {
        if (lss_status st = own_queue_stream(&b.lane_s[l])) return st;
        HIP_TRY(hipEventCreateWithFlags(&b.lane_ev[l], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&b.lane_ev2[l], hipEventDisableTiming));
    }
    return LSS_OK;
}

// Whether steps run grouped (decided once per batcher): one-graph steps, at most kGroupMax probes,
// every probe's pursuit persistent from its first pass, and all the probes' persistent workgroups
// resident at once.
lss_status group_decide(Batcher& b) {
    if (b.group >= 0) return LSS_OK;
    b.group = 0;
    const int n = (int)b.pr.size();
    if (!b.grouped || !b.single || n > lss::kGroupMax) return LSS_OK;
    std::vector<lss::GroupMember> m(n);
    for (int i = 0; i < n; ++i) {
        Probe& p = b.pr[i];
        Handle& h = *p.h;
        if (!use_persistent(h, p.W) || h.pers_first > 0 || h.prof_update) return LSS_OK;
        m[i].a = pursuit_args(h, h.B.p, h.xw.p, p.W);
        int g = 0;
        HIP_TRY(lss::persistent_grid(m[i].a, p.max_grid > 0 ? std::min(h.pers_grid, p.max_grid) : h.pers_grid, &g));
        if (g < 1) return LSS_OK;
        b.ggrid[i] = g;
        b.gmp[i] = h.prm.max_passes;
    }
    bool fits = false;
    HIP_TRY(lss::group_persistent_fits(m.data(), b.ggrid, n, &fits));
    if (!fits) return LSS_OK;
    if (lss_status st = make_lanes(b)) return st;
    b.group = 1;
    return LSS_OK;
}

// The member descriptors of the current window sets (once per set, before the first launch that
// reads them).
lss_status group_members(Batcher& b) {
    std::vector<lss::GroupMember>& m = b.ghost[b.swapped];
    if (!m.empty()) return LSS_OK;
    const int n = (int)b.pr.size();
    std::vector<lss::GroupMember> v(n);
    for (int i = 0; i < n; ++i) {
        Probe& p = b.pr[i];
        Handle& h = *p.h;
        v[i].a = pursuit_args(h, h.B.p, h.xw.p, p.W);
        v[i].f = feature_args(h, h.xw.p, p.W);
        v[i].s = h.stage;
    }
    DevBuf<lss::GroupMember>& d = b.gdev[b.swapped];
    HIP_TRY(d.alloc(n));
    HIP_TRY(hipMemcpy(d.p, v.data(), sizeof(lss::GroupMember) * n, hipMemcpyHostToDevice));
    // the same members stamping their workgroups (measurement mode)
    int total = 0;
    for (int i = 0; i < n; ++i) total += b.ggrid[i];
    if (!b.pts.p) {
        HIP_TRY(b.pts.alloc(2 * (size_t)total));
        HIP_TRY(b.pspan.alloc(2));
    }
    std::vector<lss::GroupMember> w = v;
    for (int i = 0, g0 = 0; i < n; g0 += b.ggrid[i], ++i) w[i].a.ts = b.pts.p + 2 * (size_t)g0;
    DevBuf<lss::GroupMember>& dp = b.gprof[b.swapped];
    HIP_TRY(dp.alloc(n));
    HIP_TRY(hipMemcpy(dp.p, w.data(), sizeof(lss::GroupMember) * n, hipMemcpyHostToDevice));
    m = std::move(v);
    return LSS_OK;
}

// Segment A; when no probe can need extra passes (b.single) it also holds segment C, so that a
// step is one graph.
lss_status segment_A(Batcher& b) {
    HIP_TRY(hipEventRecord(b.fork, b.main));
    for (auto& p : b.pr) {
        Handle& h = *p.h;
        HIP_TRY(hipStreamWaitEvent(h.stream, b.fork, 0));
        if (lss_status st = probe_A(p, b.cur_pre)) return st;
        if (b.single)
            if (lss_status st = probe_C(p)) return st;
        HIP_TRY(hipEventRecord(p.done, h.stream));
        HIP_TRY(hipStreamWaitEvent(b.main, p.done, 0));
    }
    return LSS_OK;
}

// Segment N (running ahead): the next batch's preprocessing of every probe into the second window
// sets, on the probes' second streams.
lss_status segment_N(Batcher& b) {
    HIP_TRY(hipEventRecord(b.fork2, b.main2));
    for (auto& p : b.pr) {
        Handle& h = *p.h;
        HIP_TRY(hipStreamWaitEvent(p.s2, b.fork2, 0));
        std::swap(h.stream, p.s2);   // the probe functions enqueue on h.stream
        const lss_status st = probe_pre(p, true);
        std::swap(h.stream, p.s2);
        if (st != LSS_OK) return st;
        HIP_TRY(hipEventRecord(p.done2, p.s2));
        HIP_TRY(hipStreamWaitEvent(b.main2, p.done2, 0));
    }
    return LSS_OK;
}

lss_status segment_C(Batcher& b) {
    HIP_TRY(hipEventRecord(b.fork, b.main));
    for (auto& p : b.pr) {
        Handle& h = *p.h;
        HIP_TRY(hipStreamWaitEvent(h.stream, b.fork, 0));
        if (lss_status st = probe_C(p)) return st;
        HIP_TRY(hipEventRecord(p.done, h.stream));
        HIP_TRY(hipStreamWaitEvent(b.main, p.done, 0));
    }
    return LSS_OK;
}

lss_status capture(Batcher& b, lss_status (*seg)(Batcher&), hipGraph_t* g, hipGraphExec_t* ge,
                   hipStream_t s = nullptr) {
    if (!s) s = b.main;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    lss_status st = seg(b);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (st != LSS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return fail(LSS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    *g = graph;
    HIP_TRY(hipGraphInstantiate(ge, graph, nullptr, nullptr, 0));
    return LSS_OK;
}

// Assembles probe p's window for batch start t0 into its pinned staging buffer: one contiguous
// copy, after the rows that repeat sample 0 when the window starts before the recording.
// Device mode: the segments that place the window's rows (samples t0 - nt .. t0 + batch + nt - 1,
// sample 0 repeated before the recording starts) into the handle's window buffer, appended to
// segs; one gather launch then copies every probe's segments.
lss_status assemble_dev(Probe& p, int64_t t0, std::vector<lss::GatherSeg>& segs, int16_t* win) {
    Handle& h = *p.h;
    const int C = h.C, nt = h.nt;
    int64_t r = 0;
    while (r < p.W && t0 - nt + r < 0) ++r;
    if (r > 0) segs.push_back({win, p.dchunks.front().first, r * C, (int64_t)C, 0});
    for (size_t k = 0; k < p.dchunks.size() && r < p.W; ++k) {
        const int64_t c0 = p.dchunks[k].second, c1 = c0 + p.dlen[k];
        const int64_t need = t0 - nt + r;   // next sample to place
        if (c1 <= need) continue;
        const int64_t from = std::max(need, c0);
        const int64_t rows = std::min(c1 - from, p.W - r);
        segs.push_back({win + r * C, p.dchunks[k].first + (from - c0) * C, rows * C, 0, 0});
        r += rows;
    }
    if (r != p.W) return fail(LSS_ERR_STATE, "device window incomplete");
    return LSS_OK;
}

void assemble(Probe& p, int64_t t0, int16_t* pin) {
    Handle& h = *p.h;
    const int C = h.C, nt = h.nt;
    int64_t c = 0;
    for (; c < p.W && t0 - nt + c < 0; ++c) std::memcpy(pin + c * C, p.buf.data() + (0 - p.base) * C, sizeof(int16_t) * C);
    const int64_t first = t0 - nt + c;
    std::memcpy(pin + c * C, p.buf.data() + (first - p.base) * C, sizeof(int16_t) * (p.W - c) * C);
}

// The second window sets, streams and events of running ahead (once per batcher and handle).
lss_status ensure_ahead(Batcher& b) {
    if (!b.main2) HIP_TRY(hipStreamCreateWithFlags(&b.main2, hipStreamNonBlocking));
    if (!b.fork2) HIP_TRY(hipEventCreateWithFlags(&b.fork2, hipEventDisableTiming));
    if (!b.joinA) HIP_TRY(hipEventCreateWithFlags(&b.joinA, hipEventDisableTiming));
    if (!b.staged) HIP_TRY(hipEventCreateWithFlags(&b.staged, hipEventDisableTiming));
    for (auto& p : b.pr) {
        Handle& h = *p.h;
        if (!h.win_alt.p) HIP_TRY(h.win_alt.alloc(h.win.n));
        if (!h.xw_alt.p) HIP_TRY(h.xw_alt.alloc(h.xw.n));
        if (!h.B_alt.p) HIP_TRY(h.B_alt.alloc(h.B.n));
        if (!p.s2) HIP_TRY(hipStreamCreateWithFlags(&p.s2, hipStreamNonBlocking));
        if (!p.done2) HIP_TRY(hipEventCreateWithFlags(&p.done2, hipEventDisableTiming));
        if (p.mode == 1 && !p.pin_alt)
            HIP_TRY(hipHostMalloc((void**)&p.pin_alt, sizeof(int16_t) * p.W * h.C, hipHostMallocDefault));
    }
    return LSS_OK;
}

bool graphs_captured(const Batcher& b) {
    if (b.gA || b.gC || b.xtail[0] || b.xtail[1]) return true;
    for (auto* e : b.execs)
        if (e) return true;
    for (int k = 0; k < 4; ++k)
        for (auto* e : b.xpre[k])
            if (e) return true;
    return false;
}

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

}  // namespace

// =================================================================================================
extern "C" {

const char* lss_last_error(void) { return g_err.c_str(); }

lss_status lss_select_device(int32_t device) {
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LSS_ERR_ARG, "lss_select_device: no such device");
    HIP_TRY(hipSetDevice(device));
    return LSS_OK;
}
int32_t lss_abi_version(void) { return LSS_ABI_VERSION; }

lss_status lss_host_alloc(size_t bytes, void** p) {
    if (!p || bytes == 0) return fail(LSS_ERR_ARG, "lss_host_alloc: bad argument");
    *p = nullptr;
    HIP_TRY(hipHostMalloc(p, bytes, hipHostMallocDefault));
    return LSS_OK;
}

lss_status lss_host_free(void* p) {
    if (p) HIP_TRY(hipHostFree(p));
    return LSS_OK;
}

lss_status lss_default_params(lss_params* p) {
    if (!p) return fail(LSS_ERR_ARG, "null params");
    std::memset(p, 0, sizeof(*p));
    p->n_chan = 384;
    p->n_templates = 0;
    p->n_pcs = 6;
    p->nt = 61;
    p->nt0min = 20;
    p->n_nearest = 10;
    p->max_window = 60122;
    p->max_passes = 50;
    p->max_spikes = 1 << 18;
    p->do_car = 1;
    p->precision = LSS_PREC_FP32;
    p->drop_negative = 1;
    p->th_learned = 8.0;
    p->fs = 30000.f;
    p->dup_bins = 7;
    return LSS_OK;
}

lss_status lss_create(const lss_params* p, const lss_artifacts* a, void** handle) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle pointer");
    Handle* h = nullptr;
    lss_status st = create_impl(p, a, &h);
    if (st == LSS_OK) *handle = h;
    return st;
}

lss_status lss_create_from_dir(const char* dir, const lss_params* p0, void** handle) {
    if (!dir || !p0 || !handle) return fail(LSS_ERR_ARG, "null argument");
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    lss_params p = *p0;
    // misc.txt key:value lines (OnlineSpikesV2.cu:435-457)
    {
        std::ifstream inf(d + "misc.txt");
        if (!inf.is_open()) return fail(LSS_ERR_IO, "cannot open " + d + "misc.txt");
        std::map<std::string, std::string> kv;
        std::string line;
        while (std::getline(inf, line)) {
            size_t pos = line.find(':');
            if (pos == std::string::npos) continue;
            kv[trim(line.substr(0, pos))] = trim(line.substr(pos + 1));
        }
        try {
            if (kv.count("nt0min")) p.nt0min = std::stoi(kv["nt0min"]);
            if (kv.count("numNearestChans")) p.n_nearest = std::stoi(kv["numNearestChans"]);
            if (kv.count("Th_learned")) p.th_learned = std::stod(kv["Th_learned"]);
            if (kv.count("duplicate_spike_bins")) p.dup_bins = std::stoi(kv["duplicate_spike_bins"]);
        } catch (...) {
            return fail(LSS_ERR_IO, "malformed misc.txt");
        }
    }
    lss::NpyArray U, wpca, wrot, drift, hp, xc, yc;
    std::string err;
    if (!lss::npy_load_f32(d + "Wall3.npy", U, err)) return fail(LSS_ERR_IO, err);
    if (!lss::npy_load_f32(d + "wPCA.npy", wpca, err)) return fail(LSS_ERR_IO, err);
    if (!lss::npy_load_f32(d + "whiteningMat.npy", wrot, err) &&
        !lss::npy_load_f32(d + "whitening_mat.npy", wrot, err))
        return fail(LSS_ERR_IO, err);
    bool have_drift = lss::npy_load_f32(d + "drift_matrix.npy", drift, err);
    if (!lss::npy_load_f32(d + "hp_filter.npy", hp, err)) return fail(LSS_ERR_IO, err);
    if (!lss::npy_load_f32(d + "xc.npy", xc, err)) return fail(LSS_ERR_IO, err);
    if (!lss::npy_load_f32(d + "yc.npy", yc, err)) return fail(LSS_ERR_IO, err);
    if (U.shape.size() != 3 || wpca.shape.size() != 2) return fail(LSS_ERR_IO, "Wall3/wPCA rank");
    if (U.shape[0] > INT32_MAX || U.shape[1] > 64 || U.shape[2] > INT32_MAX || wpca.shape[1] > 1024)
        return fail(LSS_ERR_IO, "Wall3/wPCA shape out of range");
    p.n_templates = (int)U.shape[0];
    p.n_pcs = (int)U.shape[1];
    p.n_chan = (int)U.shape[2];
    p.nt = (int)wpca.shape[1];
    if ((int)wpca.shape[0] != p.n_pcs) return fail(LSS_ERR_IO, "wPCA rows != Wall3 PCs");
    const size_t CC = (size_t)p.n_chan * p.n_chan;
    if (wrot.data.size() != CC) return fail(LSS_ERR_IO, "whitening matrix is not C x C");
    if (xc.data.size() != (size_t)p.n_chan || yc.data.size() != (size_t)p.n_chan) return fail(LSS_ERR_IO, "xc/yc size");
    std::vector<float> dm;
    if (have_drift) {
        if (drift.data.size() != CC) return fail(LSS_ERR_IO, "drift matrix is not C x C");
        // the exporter saves M.T (run_online_spikes.py:532); the matrix applied is M
        dm.resize(CC);
        for (int i = 0; i < p.n_chan; ++i)
            for (int j = 0; j < p.n_chan; ++j) dm[(size_t)i * p.n_chan + j] = drift.data[(size_t)j * p.n_chan + i];
    }
    const int64_t T = p.n_templates, K = p.n_pcs, C = p.n_chan, nC = p.n_nearest;
    // prepare_extract's tables as exported (run_online_spikes.py:533, 572-574), when present
    std::vector<int32_t> icc, iu;
    lss::NpyArray ucc;
    {
        lss::NpyArrayI64 a64;
        if (lss::npy_load_i64(d + "iCC.npy", a64, err)) {
            if (a64.shape.size() != 2 || a64.shape[0] != nC || a64.shape[1] != C)
                return fail(LSS_ERR_IO, "iCC.npy must be [numNearestChans][C]");
            icc.resize(a64.data.size());
            for (size_t i = 0; i < icc.size(); ++i) {
                if (a64.data[i] < 0 || a64.data[i] >= C) return fail(LSS_ERR_IO, "iCC.npy: channel out of range");
                icc[i] = (int32_t)a64.data[i];
            }
        }
        if (lss::npy_load_i64(d + "iU.npy", a64, err)) {
            if (a64.shape.size() != 1 || a64.shape[0] != T) return fail(LSS_ERR_IO, "iU.npy must be [T]");
            iu.resize(a64.data.size());
            for (size_t i = 0; i < iu.size(); ++i) {
                if (a64.data[i] < 0 || a64.data[i] >= C) return fail(LSS_ERR_IO, "iU.npy: channel out of range");
                iu[i] = (int32_t)a64.data[i];
            }
        }
        if (lss::npy_load_f32(d + "Ucc.npy", ucc, err)) {
            if (ucc.shape.size() != 3 || ucc.shape[0] != nC || ucc.shape[1] != T || ucc.shape[2] != K)
                return fail(LSS_ERR_IO, "Ucc.npy must be [numNearestChans][T][K]");
        }
    }
    lss_artifacts a{};
    a.U = U.data.data();
    a.wPCA = wpca.data.data();
    a.Wrot = wrot.data.data();
    a.drift = have_drift ? dm.data() : nullptr;
    a.hp_filter = hp.data.data();
    a.hp_len = (int32_t)hp.data.size();
    a.xc = xc.data.data();
    a.yc = yc.data.data();
    a.iCC = icc.empty() ? nullptr : icc.data();
    a.iU = iu.empty() ? nullptr : iu.data();
    a.Ucc = ucc.data.empty() ? nullptr : ucc.data.data();
    void* hv = nullptr;
    if (lss_status st = lss_create(&p, &a, &hv)) return st;
    // ctc.npy (the reference loads it, OnlineSpikesV2.cu:558): compared with the FP64 rebuild
    lss::NpyArray ctc;
    if (lss::npy_load_f32(d + "ctc.npy", ctc, err)) {
        Handle& h = *as_handle(hv);
        const int64_t L = 2 * (int64_t)p.nt + 1;
        if (ctc.shape.size() != 3 || ctc.shape[0] != T || ctc.shape[1] != T || ctc.shape[2] != L) {
            lss_destroy(hv);
            return fail(LSS_ERR_IO, "ctc.npy must be [T][T][2nt+1]");
        }
        std::vector<float> dev((size_t)T * L * T);   // device layout ctc[j][l][i]
        lss_status st = lss_get_tables(hv, nullptr, nullptr, dev.data(), nullptr);
        if (st != LSS_OK) {
            lss_destroy(hv);
            return st;
        }
        double md = 0.0, mx = 0.0;
        for (int64_t i = 0; i < T; ++i)
            for (int64_t j = 0; j < T; ++j)
                for (int64_t l = 0; l < L; ++l) {
                    const double r = dev[((size_t)j * L + l) * T + i];
                    const double f = ctc.data[((size_t)i * T + j) * L + l];
                    md = std::max(md, std::fabs(f - r));
                    mx = std::max(mx, std::fabs(r));
                }
        h.load.ctc_in_dir = 1;
        h.load.ctc_max_abs_diff = md;
        h.load.ctc_max_abs = mx;
    }
    *handle = hv;
    return LSS_OK;
}

lss_status lss_get_load_info(void* handle, lss_load_info* info) {
    if (!handle || !info) return fail(LSS_ERR_ARG, "null argument");
    *info = as_handle(handle)->load;
    return LSS_OK;
}

lss_status lss_destroy(void* handle) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle");
    Handle* h = as_handle(handle);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
    return LSS_OK;
}

lss_status lss_load_cluster_centroids(const char* dir, float* x, float* y, int32_t capacity, int32_t* n) {
    if (!dir || !n) return fail(LSS_ERR_ARG, "null argument");
    std::string d = dir;
    if (!d.empty() && d.back() != '/') d += '/';
    lss::NpyArray a;
    std::string err;
    if (!lss::npy_load_f32(d + "cluster_centroids.npy", a, err)) return fail(LSS_ERR_IO, err);
    if (a.data.size() % 2) return fail(LSS_ERR_IO, "cluster_centroids.npy: odd number of values (x, y pairs)");
    const int64_t m = (int64_t)a.data.size() / 2;
    if (m > INT32_MAX) return fail(LSS_ERR_IO, "cluster_centroids.npy: too many centroids");
    *n = (int32_t)m;
    if (!x || !y) return LSS_OK;   // count only
    if (capacity < m) return fail(LSS_ERR_CAPACITY, "lss_load_cluster_centroids: capacity below the centroid count");
    for (int64_t i = 0; i < m; ++i) {
        x[i] = a.data[2 * i];
        y[i] = a.data[2 * i + 1];
    }
    return LSS_OK;
}

lss_status lss_load_channel_map(const char* dir, int32_t* out, int32_t capacity, int32_t* n) {
    if (!dir || !n) return fail(LSS_ERR_ARG, "null argument");
    std::string d = dir;
    if (!d.empty() && d.back() != '/') d += '/';
    lss::NpyArray a;
    std::string err;
    if (!lss::npy_load_f32(d + "channelMap.npy", a, err)) return fail(LSS_ERR_IO, err);
    if (a.data.size() > (size_t)INT32_MAX) return fail(LSS_ERR_IO, "channelMap.npy: too many entries");
    *n = (int32_t)a.data.size();
    if (!out) return LSS_OK;   // count only
    if (capacity < *n) return fail(LSS_ERR_CAPACITY, "lss_load_channel_map: capacity below the entry count");
    for (size_t i = 0; i < a.data.size(); ++i) {
        const float v = a.data[i];
        if (!(v >= 0.f) || v != std::floor(v) || v > 16777216.f)
            return fail(LSS_ERR_IO, "channelMap.npy: entries must be non-negative integers");
        out[i] = (int32_t)v;
    }
    return LSS_OK;
}

lss_status lss_cluster_centroids_pca(void* handle, float* d_out) {
    if (!handle || !d_out) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    HIP_TRY(lss::launch_cluster_centroids(h.Ucc.p, h.T, h.nC, h.K, d_out, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

lss_status lss_get_params(void* handle, lss_params* p) {
    if (!handle || !p) return fail(LSS_ERR_ARG, "null argument");
    *p = as_handle(handle)->prm;
    return LSS_OK;
}

lss_status lss_set_stream(void* handle, void* hip_stream) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle");
    Handle* h = as_handle(handle);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->own_stream) HIP_TRY(hipStreamDestroy(h->stream));
    h->own_stream = false;
    h->stream = (hipStream_t)hip_stream;
    if (!h->stream) {
        HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->own_stream = true;
    }
    return LSS_OK;
}

lss_status lss_pad_batch(const int16_t* rec, int64_t n_samples, int32_t C, int32_t NT, int32_t nt,
                         int64_t ibatch, int16_t* out, int64_t* window_start) {
    // kilosort/io.py:584-637
    if (!rec || !out || C < 1 || NT < 1 || nt < 0 || n_samples < 1) return fail(LSS_ERR_ARG, "bad argument");
    const int64_t nb = (n_samples + NT - 1) / NT;
    if (ibatch < 0 || ibatch >= nb) return fail(LSS_ERR_ARG, "ibatch out of range");
    const int64_t Wp = (int64_t)NT + 2 * nt;
    int64_t bstart, bend;
    if (ibatch == 0) {
        bstart = 0;
        bend = std::min<int64_t>(n_samples, (int64_t)NT + nt);
    } else {
        bstart = ibatch * NT - nt;
        bend = std::min<int64_t>(n_samples, bstart + Wp);
    }
    const int64_t nsamp = bend - bstart;
    std::memset(out, 0, sizeof(int16_t) * Wp * C);
    if (ibatch == 0) {
        std::memcpy(out + (int64_t)nt * C, rec, sizeof(int16_t) * nsamp * C);
        for (int64_t w = 0; w < nt; ++w) std::memcpy(out + w * C, out + (int64_t)nt * C, sizeof(int16_t) * C);
    } else if (ibatch == nb - 1) {
        std::memcpy(out, rec + bstart * C, sizeof(int16_t) * nsamp * C);
        for (int64_t w = nsamp; w < Wp; ++w) std::memcpy(out + w * C, out + (nsamp - 1) * C, sizeof(int16_t) * C);
    } else {
        std::memcpy(out, rec + bstart * C, sizeof(int16_t) * nsamp * C);
    }
    if (window_start) *window_start = ibatch * NT - nt;
    return LSS_OK;
}

lss_status lss_sort_window(void* handle, const int16_t* window, int64_t W, int32_t on_device,
                           int64_t window_start, lss_spikes* out) {
    if (!handle || !window || !out) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (!h.alive) return fail(LSS_ERR_STATE, "handle not usable");
    if (lss_status st = check_window(h, W)) return st;
    const int16_t* d_win = window;
    HIP_TRY(hipEventRecord(h.ev[0], h.stream));
    if (!on_device) {
        HIP_TRY(hipMemcpyAsync(h.win.p, window, sizeof(int16_t) * W * h.C, hipMemcpyHostToDevice, h.stream));
        d_win = h.win.p;
    }
    if (lss_status st = run_car(h, d_win, W)) return st;
    HIP_TRY(hipEventRecord(h.ev[1], h.stream));
    if (lss_status st = run_highpass(h, h.xcw.p, h.xhp.p, W)) return st;
    HIP_TRY(hipEventRecord(h.ev[2], h.stream));
    if (lss_status st = run_whiten(h, h.xhp.p, h.xw.p, W)) return st;
    HIP_TRY(hipEventRecord(h.ev[3], h.stream));
    if (lss_status st = run_scores(h, h.xw.p, h.B.p, W, h.ev[4])) return st;
    HIP_TRY(hipEventRecord(h.ev[5], h.stream));
    if (lss_status st = run_pursuit(h, h.B.p, h.xw.p, W)) return st;
    HIP_TRY(hipEventRecord(h.ev[6], h.stream));
    if (lss_status st = run_features(h, h.xw.p, W)) return st;
    HIP_TRY(hipEventRecord(h.ev[7], h.stream));
    // counters and spikes straight into pinned host memory, one synchronisation
    if (lss_status st = ensure_stage(h)) return st;
    if (lss_status st = enqueue_readback(h)) return st;
    HIP_TRY(hipStreamSynchronize(h.stream));
    int n = 0, passes = 0;
    if (lss_status st = check_counts(h, h.hs_counters, &n, &passes)) return st;
    for (int i = 0; i < 7; ++i) HIP_TRY(hipEventElapsedTime(&h.last_ms[i], h.ev[i], h.ev[i + 1]));
    HIP_TRY(hipEventElapsedTime(&h.last_ms[7], h.ev[0], h.ev[7]));
    if (n <= h.stage_cap)
        return emit_from(h, n, passes, window_start, out, h.hs_col, h.hs_tmp, h.hs_amp, h.hs_tF, h.hs_xs, h.hs_ys);
    return emit_spikes(h, n, passes, window_start, out);
}

lss_status lss_sort_windows(void* const* handles, int32_t n_handles, const int16_t* const* windows, const int64_t* W,
                            const int64_t* window_start, int32_t on_device, int32_t n_windows, lss_spikes* outs) {
    if (!handles || n_handles < 1 || n_windows < 0 || (n_windows > 0 && (!windows || !W || !window_start || !outs)))
        return fail(LSS_ERR_ARG, "lss_sort_windows: bad argument");
    for (int j = 0; j < n_handles; ++j) {
        if (!handles[j]) return fail(LSS_ERR_ARG, "lss_sort_windows: null handle");
        for (int i = 0; i < j; ++i)
            if (handles[i] == handles[j]) return fail(LSS_ERR_ARG, "lss_sort_windows: handles must be distinct");
    }
    if (n_handles == 1 || n_windows <= 1) {
        for (int i = 0; i < n_windows; ++i)
            if (lss_status st = lss_sort_window(handles[0], windows[i], W[i], on_device, window_start[i], &outs[i]))
                return st;
        return LSS_OK;
    }
    // one host thread per handle: window i goes to handle i % n_handles, in order on that handle
    std::vector<lss_status> st(n_handles, LSS_OK);
    std::vector<std::string> msg(n_handles);
    std::vector<std::thread> th;
    th.reserve(n_handles);
    for (int j = 0; j < n_handles; ++j)
        th.emplace_back([&, j] {
            Handle& h = *as_handle(handles[j]);
            if (hipSetDevice(h.device) != hipSuccess) {
                st[j] = LSS_ERR_HIP;
                msg[j] = "hipSetDevice failed";
                return;
            }
            for (int i = j; i < n_windows; i += n_handles) {
                lss_status s = lss_sort_window(handles[j], windows[i], W[i], on_device, window_start[i], &outs[i]);
                if (s != LSS_OK) {
                    st[j] = s;
                    msg[j] = "window " + std::to_string(i) + ": " + g_err;
                    return;
                }
            }
        });
    for (auto& t : th) t.join();
    for (int j = 0; j < n_handles; ++j)
        if (st[j] != LSS_OK) return fail(st[j], msg[j]);
    return LSS_OK;
}

lss_status lss_stage_car(void* handle, const int16_t* d_window, int64_t W, float* d_out_cw) {
    if (!handle || !d_window || !d_out_cw) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (lss_status st = check_window(h, W)) return st;
    if (lss_status st = run_car(h, d_window, W)) return st;
    HIP_TRY(hipMemcpy2DAsync(d_out_cw, sizeof(float) * W, h.xcw.p, sizeof(float) * h.ldw, sizeof(float) * W, h.C,
                             hipMemcpyDeviceToDevice, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

lss_status lss_stage_filter(void* handle, const int16_t* d_window, int64_t W, float* d_out_cw) {
    if (!handle || !d_window || !d_out_cw) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (lss_status st = check_window(h, W)) return st;
    if (lss_status st = run_car(h, d_window, W)) return st;
    if (lss_status st = run_highpass(h, h.xcw.p, h.xhp.p, W)) return st;
    HIP_TRY(hipMemcpy2DAsync(d_out_cw, sizeof(float) * W, h.xhp.p, sizeof(float) * h.ldw, sizeof(float) * W, h.C,
                             hipMemcpyDeviceToDevice, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

lss_status lss_stage_highpass(void* handle, const float* d_in_cw, int64_t W, float* d_out_cw) {
    if (!handle || !d_in_cw || !d_out_cw) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (lss_status st = check_window(h, W)) return st;
    HIP_TRY(hipMemcpy2DAsync(h.xcw.p, sizeof(float) * h.ldw, d_in_cw, sizeof(float) * W, sizeof(float) * W, h.C,
                             hipMemcpyDeviceToDevice, h.stream));
    if (lss_status st = run_highpass(h, h.xcw.p, h.xhp.p, W)) return st;
    HIP_TRY(hipMemcpy2DAsync(d_out_cw, sizeof(float) * W, h.xhp.p, sizeof(float) * h.ldw, sizeof(float) * W, h.C,
                             hipMemcpyDeviceToDevice, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

// user [W][cols] <-> internal [W][ld] (device to device)
static lss_status copy_rows(Handle& h, float* dst, int64_t ldd, const float* src, int64_t lds, int64_t cols, int64_t W) {
    HIP_TRY(hipMemcpy2DAsync(dst, sizeof(float) * ldd, src, sizeof(float) * lds, sizeof(float) * cols, W,
                             hipMemcpyDeviceToDevice, h.stream));
    return LSS_OK;
}

lss_status lss_stage_whiten(void* handle, const float* d_in_cw, int64_t W, float* d_out_wc) {
    if (!handle || !d_in_cw || !d_out_wc) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (lss_status st = check_window(h, W)) return st;
    HIP_TRY(hipMemcpy2DAsync(h.xhp.p, sizeof(float) * h.ldw, d_in_cw, sizeof(float) * W, sizeof(float) * W, h.C,
                             hipMemcpyDeviceToDevice, h.stream));
    if (lss_status st = run_whiten(h, h.xhp.p, h.xw.p, W)) return st;
    if (lss_status st = copy_rows(h, d_out_wc, h.C, h.xw.p, h.ldC, h.C, W)) return st;
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

lss_status lss_stage_scores(void* handle, const float* d_x_wc, int64_t W, float* d_b_wt) {
    if (!handle || !d_x_wc || !d_b_wt) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (lss_status st = check_window(h, W)) return st;
    if (lss_status st = copy_rows(h, h.xw.p, h.ldC, d_x_wc, h.C, h.C, W)) return st;
    if (lss_status st = run_scores(h, h.xw.p, h.B.p, W)) return st;
    if (lss_status st = copy_rows(h, d_b_wt, h.T, h.B.p, h.ldT, h.T, W)) return st;
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

lss_status lss_stage_pursuit(void* handle, float* d_b_wt, float* d_x_wc, int64_t W, int64_t window_start,
                             lss_spikes* out) {
    if (!handle || !d_b_wt || !d_x_wc || !out) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (lss_status st = check_window(h, W)) return st;
    if (lss_status st = copy_rows(h, h.B.p, h.ldT, d_b_wt, h.T, h.T, W)) return st;
    if (lss_status st = copy_rows(h, h.xw.p, h.ldC, d_x_wc, h.C, h.C, W)) return st;
    if (lss_status st = run_pursuit(h, h.B.p, h.xw.p, W)) return st;
    if (lss_status st = run_features(h, h.xw.p, W)) return st;
    if (lss_status st = copy_rows(h, d_b_wt, h.T, h.B.p, h.ldT, h.T, W)) return st;
    if (lss_status st = copy_rows(h, d_x_wc, h.C, h.xw.p, h.ldC, h.C, W)) return st;
    int n = 0, passes = 0;
    if (lss_status st = read_counts(h, &n, &passes)) return st;
    return emit_spikes(h, n, passes, window_start, out);
}

// ---- template refinement ------------------------------------------------------------------------
static lss_status refine_alloc(Handle& h) {
    if (h.rf_R.p) return LSS_OK;
    HIP_TRY(h.rf_R.alloc((size_t)h.Wmax * h.ldC));
    HIP_TRY(h.rf_waves.alloc((size_t)h.T * h.nt * h.ldC));
    HIP_TRY(h.rf_dwpart.alloc((size_t)h.T * h.K * h.nt));
    HIP_TRY(h.rf_dupart.alloc((size_t)h.T * lss::refine_chunks(h.nt) * h.K * h.C));
    HIP_TRY(h.rf_part.alloc((size_t)lss::refine_blocks(h.Wmax)));
    HIP_TRY(h.rf_loss.alloc(1));
    HIP_TRY(h.rf_flag.alloc(1));
    for (auto& e : h.rf_ev) HIP_TRY(hipEventCreate(&e));
    return LSS_OK;
}

// everything lss_refine_templates needs beyond refine_alloc (done once per handle)
static lss_status refine_templates_alloc(Handle& h) {
    if (lss_status st = refine_alloc(h)) return st;
    if (h.rf_U.p) return LSS_OK;
    const size_t nU = (size_t)h.T * h.K * h.C, nW = (size_t)h.K * h.nt;
    HIP_TRY(h.rf_R2.alloc((size_t)h.Wmax * h.ldC));
    HIP_TRY(h.rf_U.alloc(nU));
    HIP_TRY(h.rf_U1.alloc(nU));
    HIP_TRY(h.rf_gU.alloc(nU));
    HIP_TRY(h.rf_gW.alloc(nW));
    HIP_TRY(h.rf_W1.alloc(nW));
    HIP_TRY(h.rf_col.alloc(h.prm.max_spikes));
    HIP_TRY(h.rf_tmp.alloc(h.prm.max_spikes));
    HIP_TRY(h.rf_amp.alloc(h.prm.max_spikes));
    HIP_TRY(h.rf_sums.alloc(256 + 4));

    return LSS_OK;
}

const char* const kRefineBusy =
    "a streaming batcher is refining this handle's templates in the background (until the step that adopts it)";

static lss_status refine_run(Handle& h, const float* d_x, int64_t W, const int32_t* d_col, const int32_t* d_tmp,
                      const float* d_amp, int64_t n, const float* d_U, const float* d_wpca, double* loss, float* d_gU,
                      float* d_gW) {
    if (!d_x || !d_U || !d_wpca || !loss || (n > 0 && (!d_col || !d_tmp || !d_amp)))
        return fail(LSS_ERR_ARG, "null argument");
    if (W < 1 || W > h.Wmax) return fail(LSS_ERR_ARG, "W must be in [1, max_window]");
    if (n < 0 || n > INT32_MAX) return fail(LSS_ERR_ARG, "bad spike count");
    if (h.rf_busy) return fail(LSS_ERR_STATE, kRefineBusy);
    if (lss_status st = refine_alloc(h)) return st;
    lss::RefineArgs a{};
    a.X = d_x; a.ldx = h.C;
    a.vec_x = (h.C % 4 == 0) && ((reinterpret_cast<uintptr_t>(d_x) & 15) == 0);
    a.col = d_col; a.tmp = d_tmp; a.amp = d_amp; a.n = (int)n;
    a.U = d_U; a.wpca = d_wpca; a.waves = h.rf_waves.p; a.R = h.rf_R.p;
    a.part = h.rf_part.p; a.loss = h.rf_loss.p; a.gU = d_gU; a.dwpart = h.rf_dwpart.p; a.dupart = h.rf_dupart.p; a.gW = d_gW;
    a.flag = h.rf_flag.p;
    a.W = W; a.ldC = h.ldC; a.T = h.T; a.C = h.C; a.K = h.K; a.nt = h.nt;
    HIP_TRY(hipEventRecord(h.rf_ev[0], h.stream));
    HIP_TRY(lss::launch_fill_i32(h.rf_flag.p, 1, 0, h.stream));
    HIP_TRY(lss::launch_waves(d_U, (int64_t)h.K * h.C, d_wpca, h.T, h.K, h.C, h.nt, h.rf_waves.p, h.ldC, h.stream));
    HIP_TRY(lss::launch_refine_residual(a, h.stream));
    HIP_TRY(hipEventRecord(h.rf_ev[1], h.stream));
    if (d_gU && d_gW) HIP_TRY(lss::launch_refine_adjoint(a, h.stream));
    HIP_TRY(hipEventRecord(h.rf_ev[2], h.stream));
    int flag = 0;
    HIP_TRY(hipMemcpyAsync(loss, h.rf_loss.p, sizeof(double), hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipMemcpyAsync(&flag, h.rf_flag.p, sizeof(int), hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    HIP_TRY(hipEventElapsedTime(&h.rf_ms[0], h.rf_ev[0], h.rf_ev[1]));
    HIP_TRY(hipEventElapsedTime(&h.rf_ms[1], h.rf_ev[1], h.rf_ev[2]));
    h.rf_ms[2] = h.rf_ms[0] + h.rf_ms[1];
    if (flag) return fail(LSS_ERR_ARG, "template id out of range in the spike list");
    return LSS_OK;
}

lss_status lss_residual_energy(void* handle, const float* d_x, int64_t W, const int32_t* d_col, const int32_t* d_tmp,
                               const float* d_amp, int64_t n, const float* d_U, const float* d_wpca, double* loss) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle");
    return refine_run(*as_handle(handle), d_x, W, d_col, d_tmp, d_amp, n, d_U, d_wpca, loss, nullptr, nullptr);
}

lss_status lss_residual_energy_grad(void* handle, const float* d_x, int64_t W, const int32_t* d_col,
                                    const int32_t* d_tmp, const float* d_amp, int64_t n, const float* d_U,
                                    const float* d_wpca, double* loss, float* d_grad_U, float* d_grad_wpca) {
    if (!handle || !d_grad_U || !d_grad_wpca) return fail(LSS_ERR_ARG, "null argument");
    return refine_run(*as_handle(handle), d_x, W, d_col, d_tmp, d_amp, n, d_U, d_wpca, loss, d_grad_U, d_grad_wpca);
}

// Loss of the last window's spike set for templates (U1 [T][K][C], wpca1) given the window's
// input x (rf_R holds it): sum (x - sum_s amp_s wave1_s)^2.  Spikes must be in rf_col/tmp/amp.
static lss_status refine_loss_at(Handle& h, hipStream_t s, int64_t W, const float* U1, const float* W1, const float* x,
                          int n, double* loss) {
    lss::RefineArgs a{};
    a.X = x; a.ldx = h.ldC; a.vec_x = 1;
    a.col = h.rf_col.p; a.tmp = h.rf_tmp.p; a.amp = h.rf_amp.p; a.n = n;
    a.U = U1; a.wpca = W1; a.waves = h.rf_waves.p; a.R = h.rf_R2.p;
    a.part = h.rf_part.p; a.loss = h.rf_loss.p; a.flag = h.rf_flag.p;
    a.W = W; a.ldC = h.ldC; a.T = h.T; a.C = h.C; a.K = h.K; a.nt = h.nt;
    HIP_TRY(lss::launch_waves(U1, (int64_t)h.K * h.C, W1, h.T, h.K, h.C, h.nt, h.rf_waves.p, h.ldC, s));
    HIP_TRY(lss::launch_refine_residual(a, s));
    HIP_TRY(hipMemcpyAsync(loss, h.rf_loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return LSS_OK;
}

// Where a refinement step writes the templates and the tables derived from them: the handle's own
// (lss_refine_templates) or a second set that the streaming batcher copies in later.
struct TemplateSet {
    float *U, *wpca, *nm, *waves, *ctc, *Ucc;
    int* iU;
};
static TemplateSet live_set(Handle& h) { return {h.U.p, h.wpca.p, h.nm.p, h.waves.p, h.ctc.p, h.Ucc.p, h.iU.p}; }

// One refinement step on stream s from the residual x [W][ldC] that the pursuit left and its n
// spikes (device col/tmp/amp, pursuit order), starting from the handle's templates; on descent the
// new U / wPCA and every table derived from them go to `out` (*adopted = true).  info as
// lss_refine_templates.  Uses the handle's refinement buffers; the handle's templates are only read.
static lss_status refine_core(Handle& h, hipStream_t s, const float* x, int64_t W, const int32_t* d_col,
                       const int32_t* d_tmp, const float* d_amp, int n, const TemplateSet& out, double* info,
                       bool* adopted) {
    *adopted = false;
    if (info)
        for (int i = 0; i < 4; ++i) info[i] = 0.0;
    if (n < 1) return LSS_OK;
    if (lss_status st = refine_templates_alloc(h)) return st;
    const int T = h.T, K = h.K, C = h.C, nt = h.nt;
    const size_t nU = (size_t)T * K * C, nW = (size_t)K * nt;
    // the window's spikes sorted by column (the residual kernel's order)
    std::vector<int> col(n), tmp(n), ord(n);
    std::vector<float> amp(n), namp(n);
    HIP_TRY(hipMemcpyAsync(col.data(), d_col, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(tmp.data(), d_tmp, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(amp.data(), d_amp, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpy2DAsync(h.rf_U.p, sizeof(float) * K * C, h.U.p, sizeof(float) * h.ldq, sizeof(float) * K * C, T,
                             hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x_, int y_) { return col[x_] < col[y_]; });
    {
        std::vector<int> c2(n), t2(n);
        std::vector<float> a2(n);
        for (int i = 0; i < n; ++i) {
            c2[i] = col[ord[i]];
            t2[i] = tmp[ord[i]];
            a2[i] = amp[ord[i]];
            namp[i] = -a2[i];
        }
        col.swap(c2);
        tmp.swap(t2);
        amp.swap(a2);
    }
    HIP_TRY(hipMemcpyAsync(h.rf_col.p, col.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(h.rf_tmp.p, tmp.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    // the window's input x = residual + sum_s amp_s wave_s (residual kernel with negated amplitudes)
    HIP_TRY(hipMemcpyAsync(h.rf_amp.p, namp.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    double* sums = h.rf_sums.p;   // [0..255] partials, [256] |g|^2
    {
        lss::RefineArgs a{};
        a.X = x; a.ldx = h.ldC; a.vec_x = 1;
        a.col = h.rf_col.p; a.tmp = h.rf_tmp.p; a.amp = h.rf_amp.p; a.n = n;
        a.U = h.rf_U.p; a.wpca = h.wpca.p; a.waves = h.rf_waves.p; a.R = h.rf_R.p;
        a.part = h.rf_part.p; a.loss = h.rf_loss.p; a.flag = h.rf_flag.p;
        a.W = W; a.ldC = h.ldC; a.T = T; a.C = C; a.K = K; a.nt = nt;
        HIP_TRY(lss::launch_waves(h.rf_U.p, (int64_t)K * C, h.wpca.p, T, K, C, nt, h.rf_waves.p, h.ldC, s));
        HIP_TRY(lss::launch_refine_residual(a, s));
        // gradient at (U0, W0) from the residual the pursuit left (R = x, which the adjoint only
        // reads); any spike order works here
        a.R = const_cast<float*>(x);
        a.amp = d_amp; a.col = d_col; a.tmp = d_tmp;
        a.gU = h.rf_gU.p; a.gW = h.rf_gW.p; a.dwpart = h.rf_dwpart.p; a.dupart = h.rf_dupart.p;
        HIP_TRY(lss::launch_refine_adjoint(a, s));
    }
    HIP_TRY(hipMemcpyAsync(h.rf_amp.p, amp.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    double g2u = 0.0, g2w = 0.0;
    HIP_TRY(lss::launch_sumsq(h.rf_gU.p, (int64_t)nU, sums, sums + 256, s));
    HIP_TRY(lss::launch_sumsq(h.rf_gW.p, (int64_t)nW, sums, sums + 257, s));
    HIP_TRY(hipMemcpyAsync(&g2u, sums + 256, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&g2w, sums + 257, sizeof(double), hipMemcpyDeviceToHost, s));
    double L0 = 0.0;
    if (lss_status st = refine_loss_at(h, s, W, h.rf_U.p, h.wpca.p, h.rf_R.p, n, &L0)) return st;   // synchronises
    const double g2 = g2u + g2w;
    if (!(g2 > 0.0) || !(L0 > 0.0)) return LSS_OK;
    // quadratic fit of L(alpha) = L(U0 - alpha gU, W0 - alpha gW) through one trial step
    auto step = [&](double al) -> lss_status {
        HIP_TRY(lss::launch_step(h.rf_U1.p, h.rf_U.p, h.rf_gU.p, al, (int64_t)nU, s));
        HIP_TRY(lss::launch_step(h.rf_W1.p, h.wpca.p, h.rf_gW.p, al, (int64_t)nW, s));
        return LSS_OK;
    };
    const double a1 = 1e-4 * L0 / g2;
    if (lss_status st = step(a1)) return st;
    double L1 = 0.0;
    if (lss_status st = refine_loss_at(h, s, W, h.rf_U1.p, h.rf_W1.p, h.rf_R.p, n, &L1)) return st;
    const double q = (L1 - L0 + a1 * g2) / (a1 * a1);
    const double al = (q > 0.0) ? std::min(g2 / (2.0 * q), 100.0 * a1) : a1;
    if (lss_status st = step(al)) return st;
    double L2 = 0.0;
    if (lss_status st = refine_loss_at(h, s, W, h.rf_U1.p, h.rf_W1.p, h.rf_R.p, n, &L2)) return st;
    if (!(L2 < L0)) return LSS_OK;   // no descent: templates unchanged
    // the step: U, wPCA, then every table derived from them
    HIP_TRY(hipMemcpy2DAsync(out.U, sizeof(float) * h.ldq, h.rf_U1.p, sizeof(float) * K * C, sizeof(float) * K * C, T,
                             hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(out.wpca, h.rf_W1.p, sizeof(float) * nW, hipMemcpyDeviceToDevice, s));
    HIP_TRY(lss::launch_extract(out.U, h.ldq, T, K, C, h.iCC.p, h.nC, out.iU, out.Ucc, s));
    HIP_TRY(lss::launch_tables(out.U, h.ldq, out.wpca, T, K, C, nt, out.nm, out.waves, h.ldC, out.ctc, h.ldT,
                               h.tables_ws.p, s));
    HIP_TRY(hipStreamSynchronize(s));
    *adopted = true;
    if (info) {
        info[0] = L0;
        info[1] = L2;
        info[2] = al;
        info[3] = std::sqrt(g2);
    }
    return LSS_OK;
}

lss_status lss_refine_templates(void* handle, double* info) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle");
    Handle& h = *as_handle(handle);
    if (info)
        for (int i = 0; i < 4; ++i) info[i] = 0.0;
    if (h.last_W < 1) return fail(LSS_ERR_STATE, "no window has been sorted yet");
    if (h.rf_busy) return fail(LSS_ERR_STATE, kRefineBusy);
    HIP_TRY(hipMemcpyAsync(h.h_counters, h.counters.p, sizeof(int) * 8, hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    const int n = std::min(h.h_counters[0] + h.h_counters[1], h.prm.max_spikes);
    bool adopted = false;
    // the handle's own templates: complete before returning (streaming steps replay graphs on
    // other streams, which must see the new tables)
    if (lss_status st = refine_core(h, h.stream, h.xw.p, h.last_W, h.sp_col.p, h.sp_tmp.p, h.sp_amp.p, n,
                                    live_set(h), info, &adopted))
        return st;
    if (adopted) ++h.pre_version;
    return LSS_OK;
}

lss_status lss_last_refine_timings(void* handle, float* ms) {
    if (!handle || !ms) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    for (int i = 0; i < 3; ++i) ms[i] = h.rf_ms[i];
    return LSS_OK;
}

lss_status lss_get_tables(void* handle, float* nm, float* waves, float* ctc, float* wmat) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle");
    Handle& h = *as_handle(handle);
    const int L = 2 * h.nt + 1;
    if (nm) HIP_TRY(hipMemcpyAsync(nm, h.nm.p, sizeof(float) * h.T, hipMemcpyDeviceToHost, h.stream));
    if (waves)
        HIP_TRY(hipMemcpy2DAsync(waves, sizeof(float) * h.C, h.waves.p, sizeof(float) * h.ldC, sizeof(float) * h.C,
                                 (size_t)h.T * h.nt, hipMemcpyDeviceToHost, h.stream));
    if (ctc)
        HIP_TRY(hipMemcpy2DAsync(ctc, sizeof(float) * h.T, h.ctc.p, sizeof(float) * h.ldT, sizeof(float) * h.T,
                                 (size_t)h.T * L, hipMemcpyDeviceToHost, h.stream));
    if (wmat)
        HIP_TRY(hipMemcpy2DAsync(wmat, sizeof(float) * h.C, h.wmat.p, sizeof(float) * h.ldm, sizeof(float) * h.C, h.C,
                                 hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

lss_status lss_set_drift(void* handle, const double* d_m) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle");
    Handle& h = *as_handle(handle);
    const int C = h.C;
    ++h.pre_version;
    if (!d_m) {
        HIP_TRY(lss::launch_f64_to_f32(h.wrot64.p, C, C, C, h.wmat.p, h.ldm, h.stream));
        HIP_TRY(hipStreamSynchronize(h.stream));
        return LSS_OK;
    }
    int S;
    int64_t kchunk;
    lss::gemm64_split(C, C, C, 0, &S, &kchunk);
    const size_t need = (size_t)(S + 1) * C * C;
    if (h.drift_ws.n < need) {
        HIP_TRY(hipStreamSynchronize(h.stream));
        if (h.drift_ws.p) HIP_TRY(hipFree(h.drift_ws.p));
        h.drift_ws.p = nullptr;
        HIP_TRY(h.drift_ws.alloc(need));
    }
    double* P = h.drift_ws.p;
    double* wm = P + (size_t)S * C * C;
    HIP_TRY(lss::launch_gemm64_kn(d_m, C, h.wrot64.p, C, P, C, C, C, S, kchunk, h.stream));
    HIP_TRY(lss::launch_reduce64(P, S, C, C, 0, 1.0, 0.0, 0.0, wm, C, h.stream));
    HIP_TRY(lss::launch_f64_to_f32(wm, C, C, C, h.wmat.p, h.ldm, h.stream));
    // done before returning: streaming steps launch on the batcher's own streams
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

lss_status lss_profile_update(void* handle, int32_t enable) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle");
    Handle& h = *as_handle(handle);
    if (enable && h.upd_ev.empty()) {
        h.upd_ev.assign(2 * (size_t)h.prm.max_passes, nullptr);
        for (auto& e : h.upd_ev) HIP_TRY(hipEventCreate(&e));
    }
    if (enable && h.upd_ts.n != (size_t)h.prm.max_passes * 2 * h.upd_grid) {
        // stamps for the current update grid (re-sized if the grid changed since)
        HIP_TRY(hipStreamSynchronize(h.stream));
        if (h.upd_ts.p) HIP_TRY(hipFree(h.upd_ts.p));
        h.upd_ts.p = nullptr;
        HIP_TRY(h.upd_ts.alloc((size_t)h.prm.max_passes * 2 * h.upd_grid));
        if (!h.upd_span.p) HIP_TRY(h.upd_span.alloc(2 * (size_t)h.prm.max_passes));
    }
    if (enable && !h.pers_ts.p) {
        HIP_TRY(h.pers_ts.alloc(2 * (size_t)kPersStamps));
        HIP_TRY(h.pers_span.alloc(2));
    }
    h.prof_update = enable != 0;
    return LSS_OK;
}

lss_status lss_last_persistent_stats(void* handle, float* gpu_ms, int32_t* spikes, int32_t* passes, int32_t* grid) {
    if (!handle || !gpu_ms || !spikes || !passes || !grid) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (!h.prof_update) return fail(LSS_ERR_STATE, "update profiling is off (lss_profile_update)");
    const bool ran = h.pers_p0 >= 0;
    *gpu_ms = ran ? h.pers_ms : 0.f;
    *spikes = ran ? h.pers_spikes : 0;
    *passes = ran ? h.pers_passes : 0;
    *grid = ran ? h.pers_used : 0;
    return LSS_OK;
}

lss_status lss_last_update_stats(void* handle, float* total_ms, int32_t* launches) {
    if (!handle || !total_ms || !launches) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    if (!h.prof_update) return fail(LSS_ERR_STATE, "update profiling is off (lss_profile_update)");
    *total_ms = h.upd_ms;
    *launches = h.upd_launches;
    return LSS_OK;
}

lss_status lss_last_update_times(void* handle, float* event_ms, float* gpu_ms, int32_t* spikes, int32_t n,
                                 int32_t* launches) {
    if (!handle || !launches || n < 0 || (n > 0 && (!event_ms || !gpu_ms || !spikes)))
        return fail(LSS_ERR_ARG, "bad argument");
    Handle& h = *as_handle(handle);
    if (!h.prof_update) return fail(LSS_ERR_STATE, "update profiling is off (lss_profile_update)");
    *launches = (int32_t)h.upd_each.size();
    for (int i = 0; i < n && i < (int)h.upd_each.size(); ++i) {
        event_ms[i] = h.upd_each[i];
        gpu_ms[i] = h.upd_gpu[i];
        spikes[i] = h.upd_spikes[i];
    }
    return LSS_OK;
}

lss_status lss_last_timings(void* handle, float* ms, int32_t n) {
    if (!handle || !ms || n < 0 || n > 8) return fail(LSS_ERR_ARG, "bad argument");
    Handle& h = *as_handle(handle);
    for (int i = 0; i < n; ++i) ms[i] = h.last_ms[i];
    return LSS_OK;
}

}  // extern "C"

// =================================================================================================
// streaming batcher and autotuner
// =================================================================================================
extern "C" {

// ---- refinement in the background ------------------------------------------------------------
// A refinement round refines probe i at the end of its i-th step (lss_stream_set_refine).  The
// step copies the window's residual and spike lists (device to device, before the next step may
// overwrite them) and a host thread runs the refinement step on the handle's own stream for it,
// writing the new templates and tables to the handle's second set; the step kRefineLag steps later
// adopts the result at its start (waiting for the thread if it has not finished), so results do
// not depend on timing, and the steps in between run with the old templates.  kRefineLag leaves the
// thread the time of a few steps (a refinement step of a 384-channel / 600-template probe takes
// ~2.5 ms beside the steps), so that the adopting step rarely waits.
constexpr int kRefineLag = 4;

static lss_status refine_buffers(Handle& h) {
    if (lss_status st = refine_templates_alloc(h)) return st;
    if (h.sh_U.p) return LSS_OK;
    HIP_TRY(hipStreamCreateWithFlags(&h.rf_stream, hipStreamNonBlocking));
    HIP_TRY(h.sn_x.alloc(h.xw.n));
    HIP_TRY(h.sn_col.alloc(h.prm.max_spikes));
    HIP_TRY(h.sn_tmp.alloc(h.prm.max_spikes));
    HIP_TRY(h.sn_amp.alloc(h.prm.max_spikes));
    HIP_TRY(h.sh_U.alloc(h.U.n));
    HIP_TRY(h.sh_wpca.alloc(h.wpca.n));
    HIP_TRY(h.sh_nm.alloc(h.nm.n));
    HIP_TRY(h.sh_waves.alloc(h.waves.n));
    HIP_TRY(h.sh_ctc.alloc(h.ctc.n));
    HIP_TRY(h.sh_Ucc.alloc(h.Ucc.n));
    HIP_TRY(h.sh_iU.alloc(h.iU.n));
    return LSS_OK;
}

// Waits for probe p's job (if not done yet) and keeps its info.
static lss_status refine_join(Batcher& b, Probe& p) {
    if (p.rjob_adopt < 0 || p.rjob_joined) return LSS_OK;
    p.rjob.join();
    p.rjob_joined = true;
    p.h->rf_busy = false;
    if (p.rjob_st != LSS_OK) return fail(p.rjob_st, "template refinement: " + p.rjob_err);
    b.refine_info.resize(b.pr.size());
    b.refine_info[&p - b.pr.data()] = p.rjob_info;
    return LSS_OK;
}

// Adopts probe p's refinement: its templates and tables replace the handle's, in stream order
// before the step's work (the copies run on b.main).
static lss_status refine_adopt(Batcher& b, Probe& p) {
    if (p.rjob_adopt < 0) return LSS_OK;
    if (lss_status st = refine_join(b, p)) return st;
    p.rjob_adopt = -1;
    p.rjob_joined = false;
    if (!p.rjob_adopted) return LSS_OK;
    Handle& h = *p.h;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, b.main);
    };
    HIP_TRY(cp(h.U.p, h.sh_U.p, sizeof(float) * h.U.n));
    HIP_TRY(cp(h.wpca.p, h.sh_wpca.p, sizeof(float) * h.wpca.n));
    HIP_TRY(cp(h.nm.p, h.sh_nm.p, sizeof(float) * h.nm.n));
    HIP_TRY(cp(h.waves.p, h.sh_waves.p, sizeof(float) * h.waves.n));
    HIP_TRY(cp(h.ctc.p, h.sh_ctc.p, sizeof(float) * h.ctc.n));
    HIP_TRY(cp(h.Ucc.p, h.sh_Ucc.p, sizeof(float) * h.Ucc.n));
    HIP_TRY(cp(h.iU.p, h.sh_iU.p, sizeof(int) * h.iU.n));
    ++h.pre_version;
    return LSS_OK;
}

// Starts probe p's refinement on the window of the step that just ended (its residual and spikes).
static lss_status refine_start(Batcher& b, Probe& p) {
    if (lss_status st = refine_adopt(b, p)) return st;   // one job per probe at a time
    Handle& h = *p.h;
    if (lss_status st = refine_buffers(h)) return st;
    if (!p.rsnap) HIP_TRY(hipEventCreateWithFlags(&p.rsnap, hipEventDisableTiming));
    const int n = std::min(h.hs_counters[0] + h.hs_counters[1], h.prm.max_spikes);
    HIP_TRY(hipMemcpyAsync(h.sn_x.p, h.xw.p, sizeof(float) * p.W * h.ldC, hipMemcpyDeviceToDevice, h.rf_stream));
    HIP_TRY(hipMemcpyAsync(h.sn_col.p, h.sp_col.p, sizeof(int) * n, hipMemcpyDeviceToDevice, h.rf_stream));
    HIP_TRY(hipMemcpyAsync(h.sn_tmp.p, h.sp_tmp.p, sizeof(int) * n, hipMemcpyDeviceToDevice, h.rf_stream));
    HIP_TRY(hipMemcpyAsync(h.sn_amp.p, h.sp_amp.p, sizeof(float) * n, hipMemcpyDeviceToDevice, h.rf_stream));
    HIP_TRY(hipEventRecord(p.rsnap, h.rf_stream));
    HIP_TRY(hipStreamWaitEvent(b.main, p.rsnap, 0));   // the next step writes them only after the copies
    if (b.main2) HIP_TRY(hipStreamWaitEvent(b.main2, p.rsnap, 0));
    p.rjob_adopt = b.batches + kRefineLag - 1;   // b.batches already counts the step that just ended
    p.rjob_joined = false;
    p.rjob_st = LSS_OK;
    p.rjob_adopted = false;
    p.rjob_info = {};
    const int64_t W = p.W;
    Handle* hp = &h;
    Probe* pp = &p;
    h.rf_busy = true;
    p.rjob = std::thread([hp, pp, n, W] {
        Handle& hh = *hp;
        if (hipSetDevice(hh.device) != hipSuccess) {
            pp->rjob_st = LSS_ERR_HIP;
            pp->rjob_err = "hipSetDevice failed";
            return;
        }
        const TemplateSet out{hh.sh_U.p, hh.sh_wpca.p, hh.sh_nm.p, hh.sh_waves.p, hh.sh_ctc.p, hh.sh_Ucc.p,
                              hh.sh_iU.p};
        bool adopted = false;
        pp->rjob_st = refine_core(hh, hh.rf_stream, hh.sn_x.p, W, hh.sn_col.p, hh.sn_tmp.p, hh.sn_amp.p, n, out,
                                  pp->rjob_info.data(), &adopted);
        if (pp->rjob_st != LSS_OK) pp->rjob_err = g_err;
        pp->rjob_adopted = adopted;
    });
    return LSS_OK;
}

lss_status lss_stream_create(void* const* handles, int32_t n_probes, int32_t batch, int32_t use_graph,
                             void** stream_out) {
    if (!handles || n_probes < 1 || batch < 1 || !stream_out) return fail(LSS_ERR_ARG, "bad argument");
    std::unique_ptr<Batcher> b(new Batcher());
    b->batch = batch;
    b->use_graph = use_graph != 0;
    if (lss_status st = own_queue_stream(&b->main)) return st;   // no other stream's work queued before a step's
    HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    b->pr.resize(n_probes);
    for (int i = 0; i < n_probes; ++i) {
        Probe& p = b->pr[i];
        if (!handles[i]) return fail(LSS_ERR_ARG, "null handle");
        p.h = as_handle(handles[i]);
        Handle& h = *p.h;
        for (int j = 0; j < i; ++j)
            if (b->pr[j].h == p.h) return fail(LSS_ERR_ARG, "a handle appears twice");
        p.W = (int64_t)batch + 2 * h.nt;
        if (p.W > h.Wmax) return fail(LSS_ERR_ARG, "batch + 2*nt exceeds a handle's max_window");
        if (h.prof_update) return fail(LSS_ERR_STATE, "turn update profiling off before streaming");
        if (lss_status st = ensure_stage(h)) return st;
        if (lss_status st = get_spectrum(h, p.W, true, &p.spec)) return st;
        HIP_TRY(hipStreamSynchronize(h.stream));   // in place before any replay on other streams
        HIP_TRY(hipHostMalloc((void**)&p.pin, sizeof(int16_t) * p.W * h.C, hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&p.done, hipEventDisableTiming));
        p.P1 = (h.tuned_passes > 0) ? std::min(h.tuned_passes, h.prm.max_passes) : h.prm.max_passes;
        if (use_persistent(h, p.W)) {
            p.P1 = h.prm.max_passes;
            HIP_TRY(lss::persistent_max_grid(h.T, p.W, &p.max_grid));
        }
    }
    // the probes' persistent kernels run concurrently: probe i may use the fraction
    // f_i = sqrt(T_i C_i) / sum_j sqrt(T_j C_j) of the workgroups the device holds of its kernel, so
    // that the fractions of all probes add up to at most one device
    double wsum = 0.0;
    for (auto& p : b->pr) wsum += std::sqrt((double)p.h->T * p.h->C);
    b->single = true;
    for (auto& p : b->pr) {
        if (p.max_grid > 0)
            p.max_grid = std::max(1, (int)std::floor(p.max_grid * std::sqrt((double)p.h->T * p.h->C) / wsum));
        if (p.P1 < p.h->prm.max_passes) b->single = false;
    }
    *stream_out = b.release();
    return LSS_OK;
}

lss_status lss_stream_destroy(void* stream) {
    if (!stream) return fail(LSS_ERR_ARG, "null stream");
    Batcher* b = static_cast<Batcher*>(stream);
    (void)hipStreamSynchronize(b->main);
    delete b;
    return LSS_OK;
}

lss_status lss_stream_push(void* stream, int32_t probe, const int16_t* samples, int64_t n) {
    if (!stream || (n > 0 && !samples) || n < 0) return fail(LSS_ERR_ARG, "bad argument");
    Batcher& b = *static_cast<Batcher*>(stream);
    if (probe < 0 || probe >= (int32_t)b.pr.size()) return fail(LSS_ERR_ARG, "probe out of range");
    Probe& p = b.pr[probe];
    if (p.mode == 2) return fail(LSS_ERR_STATE, "this probe receives device samples");
    p.mode = 1;
    const int C = p.h->C;
    p.buf.insert(p.buf.end(), samples, samples + n * C);
    p.recv += n;
    return LSS_OK;
}

lss_status lss_stream_push_device(void* stream, int32_t probe, const int16_t* d_samples, int64_t n) {
    if (!stream || (n > 0 && !d_samples) || n < 0) return fail(LSS_ERR_ARG, "bad argument");
    Batcher& b = *static_cast<Batcher*>(stream);
    if (probe < 0 || probe >= (int32_t)b.pr.size()) return fail(LSS_ERR_ARG, "probe out of range");
    Probe& p = b.pr[probe];
    if (p.mode == 1) return fail(LSS_ERR_STATE, "this probe receives host samples");
    if (graphs_captured(b) && p.mode == 0) return fail(LSS_ERR_STATE, "the step graph was captured for host samples");
    p.mode = 2;
    if (n == 0) return LSS_OK;
    p.dchunks.emplace_back(d_samples, p.recv);
    p.dlen.push_back(n);
    p.recv += n;
    return LSS_OK;
}

// The window sets preprocessed ahead become the current ones (and the current ones the second).
static void swap_sets(Batcher& b) {
    for (auto& p : b.pr) {
        Handle& h = *p.h;
        std::swap(h.win.p, h.win_alt.p);
        std::swap(h.xw.p, h.xw_alt.p);
        std::swap(h.B.p, h.B_alt.p);
        std::swap(p.pin, p.pin_alt);
    }
    b.swapped ^= 1;
}

lss_status lss_stream_step(void* stream, lss_spikes* outs, int32_t* ready) {
    if (!stream || !outs || !ready) return fail(LSS_ERR_ARG, "null argument");
    Batcher& b = *static_cast<Batcher*>(stream);
    *ready = 0;
    for (auto& p : b.pr)
        if (p.recv < b.t0 + b.batch + p.h->nt) return LSS_OK;   // not enough samples yet
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point t) { return std::chrono::duration<float, std::milli>(clk::now() - t).count(); };
    auto t_start = clk::now();
    // refinements due at this step replace their probes' templates first
    for (auto& p : b.pr)
        if (p.rjob_adopt >= 0 && p.rjob_adopt <= b.batches)
            if (lss_status st = refine_adopt(b, p)) return st;
    // this batch preprocessed ahead by the last step: its sets become the current ones, unless an
    // input of some probe's preprocessing changed since (then every probe preprocesses it again)
    bool have = b.ahead;
    for (auto& p : b.pr)
        if (p.h->pre_version != p.ahead_version) have = false;
    b.tail_inflight = false;
    if (b.early) {
        // the last step launched this batch's tail (sets already swapped); it stands unless a handle
        // was used or an input of the preprocessing changed since
        bool ok = true;
        for (auto& p : b.pr)
            if (p.h->touched || p.h->pre_version != p.ahead_version) ok = false;
        if (!ok) HIP_TRY(hipStreamSynchronize(b.main));
        b.tail_inflight = ok;
        have = ok;
        b.early = false;
    } else if (have) {
        swap_sets(b);
    }
    for (auto& p : b.pr) p.h->touched = false;
    b.ahead = false;
    // the next batch ahead: one-graph steps, its samples here, and no refinement adopted at the next
    // step's start (that changes the templates its scores need)
    bool next = b.single && b.run_ahead;
    for (auto& p : b.pr)
        if (p.rjob_adopt == b.batches + 1) next = false;
    for (auto& p : b.pr)
        if (p.recv < b.t0 + 2 * b.batch + p.h->nt) next = false;
    if (next)
        if (lss_status st = ensure_ahead(b)) return st;
    b.cur_pre = !have;
    b.next_pre = next;
    std::vector<lss::GatherSeg> segs;
    for (auto& p : b.pr) {
        Handle& h = *p.h;
        for (int k = 0; k < 2; ++k) {
            if (!(k == 0 ? b.cur_pre : b.next_pre)) continue;
            const int64_t t = b.t0 + k * b.batch;
            if (p.mode == 1) assemble(p, t, k == 0 ? p.pin : p.pin_alt);
            else if (lss_status st = assemble_dev(p, t, segs, k == 0 ? h.win.p : h.win_alt.p)) return st;
        }
    }
    if (!segs.empty())
        HIP_TRY(lss::launch_gather(segs.data(), (int)segs.size(), b.tail_inflight ? b.lane_s[0] : b.main));
    if (b.single) {
        if (lss_status st = group_decide(b)) return st;
        if (b.group == 1)
            if (lss_status st = group_members(b)) return st;
    }
    if (b.next_pre && b.group != 1) {
        // segment N reads the windows just gathered and the templates a refinement adopted at the
        // start of this step, both written on b.main
        HIP_TRY(hipEventRecord(b.staged, b.main));
        HIP_TRY(hipStreamWaitEvent(b.main2, b.staged, 0));
    }
    b.part_ms[0] = ms_since(t_start);
    auto t1 = clk::now();
    // segment A
    if (b.group == 1) {
        if (lss_status st = grouped_step(b)) return st;
        for (auto& p : b.pr)
            if (b.early_ev && p.h->pending == b.early_ev) p.h->pending = nullptr;
    } else if (b.use_graph && b.single) {
        const int key = (b.cur_pre ? 1 : 0) | (b.swapped << 1);
        if (!b.execs[key])
            if (lss_status st = capture(b, segment_A, &b.graphs[key], &b.execs[key])) return st;
        HIP_TRY(hipGraphLaunch(b.execs[key], b.main));
    } else if (b.use_graph) {
        if (!b.gA)
            if (lss_status st = capture(b, segment_A, &b.graphA, &b.gA)) return st;
        HIP_TRY(hipGraphLaunch(b.gA, b.main));
    } else if (lss_status st = segment_A(b)) {
        return st;
    }
    if (b.next_pre && b.group != 1) {
        // segment N beside segment A; after it when this step also preprocesses its own batch
        // (the two share the handles' intermediate buffers)
        if (b.cur_pre) {
            HIP_TRY(hipEventRecord(b.joinA, b.main));
            HIP_TRY(hipStreamWaitEvent(b.main2, b.joinA, 0));
        }
        if (b.use_graph) {
            if (!b.execsN[b.swapped])
                if (lss_status st = capture(b, segment_N, &b.graphsN[b.swapped], &b.execsN[b.swapped], b.main2))
                    return st;
            HIP_TRY(hipGraphLaunch(b.execsN[b.swapped], b.main2));
        } else if (lss_status st = segment_N(b)) {
            return st;
        }
        HIP_TRY(hipStreamSynchronize(b.main2));
    }
    HIP_TRY(hipStreamSynchronize(b.main));
    b.part_ms[1] = ms_since(t1);
    t1 = clk::now();
    // passes past P1 for probes that have not converged
    for (auto& p : b.pr) {
        if (b.single) break;
        Handle& h = *p.h;
        if (p.P1 < h.prm.max_passes && h.hs_counters[3] == 0) {
            const lss::PursuitArgs a = pursuit_args(h, h.B.p, h.xw.p, p.W);
            if (lss_status st = enqueue_passes(h, a, p.P1, h.prm.max_passes)) return st;
            HIP_TRY(hipEventRecord(p.done, h.stream));
            HIP_TRY(hipStreamWaitEvent(b.main, p.done, 0));
        }
    }
    b.part_ms[2] = ms_since(t1);
    t1 = clk::now();
    // segment C (already in segment A when b.single)
    if (!b.single) {
        if (b.use_graph) {
            if (!b.gC)
                if (lss_status st = capture(b, segment_C, &b.graphC, &b.gC)) return st;
            HIP_TRY(hipGraphLaunch(b.gC, b.main));
        } else if (lss_status st = segment_C(b)) {
            return st;
        }
        HIP_TRY(hipStreamSynchronize(b.main));
    }
    b.part_ms[3] = ms_since(t1);
    t1 = clk::now();
    for (size_t i = 0; i < b.pr.size(); ++i) {
        Probe& p = b.pr[i];
        Handle& h = *p.h;
        const int* c = h.hs_counters;
        if (c[7]) return fail(LSS_ERR_STATE, kBarrierTimeout);
        if (c[2]) return fail(LSS_ERR_CAPACITY, "spike capacity (max_spikes) exceeded");
        if (c[5])
            return fail(LSS_ERR_CAPACITY, "more than 1024 spikes of a single pass (tied peaks) reach one 16-column "
                                          "block of the residual; the residual is incomplete");
        const int n = c[0] + c[1];
        const int64_t ws = b.t0 - h.nt;
        lss_status st = (n <= h.stage_cap)
                            ? emit_from(h, n, c[4], ws, &outs[i], h.hs_col, h.hs_tmp, h.hs_amp, h.hs_tF, h.hs_xs, h.hs_ys)
                            : emit_spikes(h, n, c[4], ws, &outs[i]);
        if (st != LSS_OK) return st;
    }
    b.part_ms[4] = ms_since(t1);
    if (next) {
        b.ahead = true;
        ++b.ran_ahead;
        for (auto& p : b.pr) p.ahead_version = p.h->pre_version;
    }
    ++b.batches;
    bool refined = false;
    if (b.refine_every > 0) {
        // a round starts every refine_every steps (one still under way carries on); in a round,
        // step i refines probe i, so the step that refines pays for one probe's refinement only
        if (b.batches % b.refine_every == 0 && b.refine_next < 0) b.refine_next = 0;
        if (b.refine_next >= 0) {
            const int k = b.refine_next;
            b.refine_info.resize(b.pr.size());
            if (lss_status st = refine_start(b, b.pr[k])) return st;
            refined = true;
            b.refine_next = (k + 1 < (int)b.pr.size()) ? k + 1 : -1;
        }
    }
    b.part_ms[5] = ms_since(t_start);
    b.last_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    b.t0 += b.batch;
    for (auto& p : b.pr) {
        // keep samples from t0 - nt on
        const int64_t keep_from = std::max<int64_t>(0, b.t0 - p.h->nt);
        const int64_t drop = keep_from - p.base;
        if (p.mode == 2) {
            while (!p.dchunks.empty() && p.dchunks.front().second + p.dlen.front() <= keep_from) {
                p.dchunks.pop_front();
                p.dlen.pop_front();
            }
            p.base = keep_from;
        } else if (drop > 0) {
            p.buf.erase(p.buf.begin(), p.buf.begin() + drop * p.h->C);
            p.base = keep_from;
        }
    }
    // pipelined: the next batch, preprocessed by this step, starts its pursuit now (not after a step
    // that started a refinement, whose snapshot the next preprocessing must follow, nor before a step
    // that adopts one)
    if (b.pipelined && b.group == 1 && b.ahead && !b.profile && !refined) {
        bool ok = true;
        for (auto& p : b.pr)
            if (p.rjob_adopt >= 0 && p.rjob_adopt <= b.batches) ok = false;
        if (ok) {
            swap_sets(b);
            b.ahead = false;
            if (lss_status st = group_members(b)) return st;
            if (lss_status st = launch_tail(b)) return st;
            if (!b.early_ev) HIP_TRY(hipEventCreateWithFlags(&b.early_ev, hipEventDisableTiming));
            HIP_TRY(hipEventRecord(b.early_ev, b.main));
            for (auto& p : b.pr) p.h->pending = b.early_ev;
            b.early = true;
            ++b.pipelined_steps;
        }
    }
    *ready = 1;
    return LSS_OK;
}

lss_status lss_stream_stats(void* stream, float* last_ms, int64_t* batches) {
    if (!stream || !last_ms || !batches) return fail(LSS_ERR_ARG, "null argument");
    Batcher& b = *static_cast<Batcher*>(stream);
    *last_ms = b.last_ms;
    *batches = b.batches;
    return LSS_OK;
}

lss_status lss_stream_set_refine(void* stream, int32_t every_batches) {
    if (!stream || every_batches < 0) return fail(LSS_ERR_ARG, "bad argument");
    Batcher& b = *static_cast<Batcher*>(stream);
    // the refinement buffers are allocated now, not in the step that first refines
    if (every_batches > 0)
        for (auto& p : b.pr)
            if (lss_status st = refine_buffers(*p.h)) return st;
    b.refine_every = every_batches;
    return LSS_OK;
}

lss_status lss_stream_set_run_ahead(void* stream, int32_t on) {
    if (!stream) return fail(LSS_ERR_ARG, "null stream");
    static_cast<Batcher*>(stream)->run_ahead = on != 0;
    return LSS_OK;
}

lss_status lss_stream_set_grouped(void* stream, int32_t on) {
    if (!stream) return fail(LSS_ERR_ARG, "null stream");
    Batcher& b = *static_cast<Batcher*>(stream);
    if (b.batches > 0 || graphs_captured(b)) return fail(LSS_ERR_STATE, "set grouping before the first step");
    b.grouped = on != 0;
    b.group = -1;
    return LSS_OK;
}

lss_status lss_stream_grouped(void* stream, int32_t* grouped) {
    if (!stream || !grouped) return fail(LSS_ERR_ARG, "null argument");
    *grouped = static_cast<Batcher*>(stream)->group == 1 ? 1 : 0;
    return LSS_OK;
}

lss_status lss_stream_profile(void* stream, int32_t on) {
    if (!stream) return fail(LSS_ERR_ARG, "null stream");
    static_cast<Batcher*>(stream)->profile = on != 0;
    return LSS_OK;
}

lss_status lss_stream_last_profile(void* stream, double* gpu_us, int32_t* spikes, int32_t* passes, int32_t* grid) {
    if (!stream || !gpu_us || !spikes || !passes || !grid) return fail(LSS_ERR_ARG, "null argument");
    Batcher& b = *static_cast<Batcher*>(stream);
    if (!b.profile) return fail(LSS_ERR_STATE, "stream measurement is off (lss_stream_profile)");
    if (b.group != 1) return fail(LSS_ERR_STATE, "the stream's steps do not run grouped");
    *gpu_us = b.prof_us;
    for (size_t i = 0; i < b.pr.size(); ++i) {
        const int* c = b.pr[i].h->hs_counters;
        spikes[i] = c[0] + c[1];
        passes[i] = c[4];
        grid[i] = b.ggrid[i];
    }
    return LSS_OK;
}

lss_status lss_stream_set_pipelined(void* stream, int32_t on) {
    if (!stream) return fail(LSS_ERR_ARG, "null stream");
    static_cast<Batcher*>(stream)->pipelined = on != 0;
    return LSS_OK;
}

lss_status lss_stream_pipelined(void* stream, int64_t* steps) {
    if (!stream || !steps) return fail(LSS_ERR_ARG, "null argument");
    *steps = static_cast<Batcher*>(stream)->pipelined_steps;
    return LSS_OK;
}

lss_status lss_stream_ran_ahead(void* stream, int64_t* steps) {
    if (!stream || !steps) return fail(LSS_ERR_ARG, "null argument");
    *steps = static_cast<Batcher*>(stream)->ran_ahead;
    return LSS_OK;
}

lss_status lss_stream_last_refine(void* stream, int32_t probe, double* info) {
    if (!stream || !info) return fail(LSS_ERR_ARG, "null argument");
    Batcher& b = *static_cast<Batcher*>(stream);
    if (probe < 0 || probe >= (int32_t)b.pr.size()) return fail(LSS_ERR_ARG, "probe out of range");
    if (lss_status st = refine_join(b, b.pr[probe])) return st;   // a step still under way: its result
    for (int i = 0; i < 4; ++i) info[i] = (probe < (int32_t)b.refine_info.size()) ? b.refine_info[probe][i] : 0.0;
    return LSS_OK;
}

lss_status lss_stream_breakdown(void* stream, float* ms) {
    if (!stream || !ms) return fail(LSS_ERR_ARG, "null argument");
    Batcher& b = *static_cast<Batcher*>(stream);
    for (int i = 0; i < 6; ++i) ms[i] = b.part_ms[i];
    return LSS_OK;
}

lss_status lss_set_tuning(void* handle, int32_t update_grid, int32_t first_passes) {
    if (!handle || update_grid < 1 || first_passes < 0) return fail(LSS_ERR_ARG, "bad argument");
    Handle& h = *as_handle(handle);
    h.upd_grid = update_grid;
    h.tuned_passes = first_passes;
    return LSS_OK;
}

lss_status lss_get_tuning(void* handle, int32_t* update_grid, int32_t* first_passes) {
    if (!handle || !update_grid || !first_passes) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    *update_grid = h.upd_grid;
    *first_passes = h.tuned_passes;
    return LSS_OK;
}

lss_status lss_set_persistent(void* handle, int32_t grid, int32_t first_passes) {
    if (!handle || grid < 0 || first_passes < 0) return fail(LSS_ERR_ARG, "bad argument");
    Handle& h = *as_handle(handle);
    h.pers_grid = grid;
    h.pers_first = first_passes;
    return LSS_OK;
}

lss_status lss_get_persistent(void* handle, int64_t W, int32_t* grid, int32_t* first_passes, int32_t* max_grid) {
    if (!handle || !grid || !first_passes || !max_grid) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    *grid = h.pers_grid;
    *first_passes = h.pers_first;
    *max_grid = 0;
    if (W >= 1 && W <= lss::kPersistMaxW) HIP_TRY(lss::persistent_max_grid(h.T, W, max_grid));
    return LSS_OK;
}

lss_status lss_autotune_persistent(void* handle, const int16_t* d_window, int64_t W, int32_t reps, int32_t max_grid,
                                   float* ms, int32_t* best_grid, int32_t* best_first) {
    if (!handle || !d_window || reps < 1 || max_grid < 0 || !best_grid || !best_first)
        return fail(LSS_ERR_ARG, "bad argument");
    if (W > lss::kPersistMaxW) return fail(LSS_ERR_ARG, "window longer than LSS_PERSIST_MAX_WINDOW");
    Handle& h = *as_handle(handle);
    int cap = 0;
    HIP_TRY(lss::persistent_max_grid(h.T, W, &cap));
    if (max_grid > 0) cap = std::min(cap, max_grid);
    static const int kGrid[] = {0, 64, 128, 256, 384};
    // leading per-pass rounds: none, 2, 4, or the tuned first-segment pass count (lss_autotune),
    // after which the persistent kernel usually finds the pursuit converged and exits at once
    const int kFirst[] = {0, 2, 4, (h.tuned_passes > 0 && h.tuned_passes < h.prm.max_passes) ? h.tuned_passes : -1};
    lss_spikes out{};
    out.capacity = h.prm.max_spikes;
    float best = 1e30f;
    const int g0 = h.pers_grid, f0 = h.pers_first;
    int bg = g0, bf = f0;
    for (int gi = 0; gi < 5; ++gi)
        for (int fi = 0; fi < 4; ++fi) {
            if (ms) ms[gi * 4 + fi] = -1.f;
            // candidates above the limit run at the limit once
            const int g = std::min(kGrid[gi], cap);
            if ((gi > 0 && g <= std::min(kGrid[gi - 1], cap)) || (g == 0 && fi > 0) || kFirst[fi] < 0 ||
                kFirst[fi] >= h.prm.max_passes)
                continue;
            h.pers_grid = g;
            h.pers_first = kFirst[fi];
            float t = 1e30f;
            for (int r = 0; r <= reps; ++r) {
                if (lss_status st = lss_sort_window(handle, d_window, W, 1, 0, &out)) {
                    h.pers_grid = g0;
                    h.pers_first = f0;
                    return st;
                }
                if (r > 0) t = std::min(t, h.last_ms[7]);
            }
            if (ms) ms[gi * 4 + fi] = t;
            if (t < best) {
                best = t;
                bg = g;
                bf = kFirst[fi];
            }
        }
    h.pers_grid = bg;
    h.pers_first = bf;
    *best_grid = bg;
    *best_first = bf;
    return LSS_OK;
}

lss_status lss_set_tiles(void* handle, int32_t whiten, int32_t project, int32_t scores) {
    if (!handle) return fail(LSS_ERR_ARG, "null handle");
    for (int t : {whiten, project, scores})
        if (t < 0 || t >= lss::kGemmTiles) return fail(LSS_ERR_ARG, "tile shape out of range");
    Handle& h = *as_handle(handle);
    h.tile_whiten = whiten;
    h.proj_tile = project;
    h.tile_scores = scores;
    return LSS_OK;
}

lss_status lss_get_tiles(void* handle, int32_t* tiles) {
    if (!handle || !tiles) return fail(LSS_ERR_ARG, "null argument");
    Handle& h = *as_handle(handle);
    tiles[0] = h.tile_whiten;
    tiles[1] = h.proj_tile;
    tiles[2] = h.tile_scores;
    return LSS_OK;
}

lss_status lss_autotune_tiles(void* handle, const int16_t* d_window, int64_t W, int32_t reps, float* ms,
                              int32_t* best) {
    if (!handle || !d_window || reps < 1 || !best) return fail(LSS_ERR_ARG, "bad argument");
    Handle& h = *as_handle(handle);
    if (!h.alive) return fail(LSS_ERR_STATE, "handle not usable");
    if (lss_status st = check_window(h, W)) return st;
    // the window's filtered data as the stages below see it
    if (lss_status st = run_car(h, d_window, W)) return st;
    if (lss_status st = run_highpass(h, h.xcw.p, h.xhp.p, W)) return st;
    const bool x3 = h.prm.precision == LSS_PREC_BF16X3 || h.prm.precision == LSS_PREC_FP16X3;
    int* knob[3] = {&h.tile_whiten, &h.proj_tile, &h.tile_scores};
    // stage 0: whitening GEMM, 1: projection, 2: score GEMM; each on the previous stage's output
    auto run_stage = [&](int stage) -> lss_status {
        if (stage == 0) return run_whiten(h, h.xhp.p, h.xw.p, W);
        const int KC = h.K * h.C;
        if (stage == 1) {
            if (x3)
                HIP_TRY(lss::launch_project_x3(h.prm.precision, h.xw.p, h.ldC, W, h.C, h.wpca.p, h.K, h.nt, h.xp.p,
                                               h.ldq, h.stream, h.proj_tile));
            else
                HIP_TRY(lss::launch_project(h.xw.p, h.ldC, W, h.C, h.wpca.p, h.K, h.nt, h.xp.p, h.ldq, h.stream,
                                            h.proj_tile));
            return LSS_OK;
        }
        if (h.prm.precision == LSS_PREC_FP32)
            HIP_TRY(lss::launch_gemm_f32(h.xp.p, h.ldq, 0, h.U.p, h.ldq, h.B.p, h.ldT, (int)W, h.T, KC, h.stream,
                                         h.tile_scores));
        else
            HIP_TRY(lss::launch_gemm_lp(h.prm.precision, h.xp.p, h.ldq, 0, h.U.p, h.ldq, h.B.p, h.ldT, (int)W, h.T, KC,
                                        h.stream, h.tile_scores));
        return LSS_OK;
    };
    for (int stage = 0; stage < 3; ++stage) {
        const int keep = *knob[stage];
        float bt = 1e30f;
        int bi = keep;
        for (int t = 0; t < lss::kGemmTiles; ++t) {
            *knob[stage] = t;
            float tm = 1e30f;
            for (int r = 0; r <= reps; ++r) {   // the first run warms up
                HIP_TRY(hipEventRecord(h.ev[0], h.stream));
                if (lss_status st = run_stage(stage)) {
                    *knob[stage] = keep;
                    return st;
                }
                HIP_TRY(hipEventRecord(h.ev[1], h.stream));
                HIP_TRY(hipEventSynchronize(h.ev[1]));
                float e = 0.f;
                HIP_TRY(hipEventElapsedTime(&e, h.ev[0], h.ev[1]));
                if (r > 0) tm = std::min(tm, e);
            }
            if (ms) ms[stage * lss::kGemmTiles + t] = tm;
            if (tm < bt) {
                bt = tm;
                bi = t;
            }
        }
        *knob[stage] = bi;
        best[stage] = bi;
        if (lss_status st = run_stage(stage)) return st;   // the next stage's input
    }
    HIP_TRY(hipStreamSynchronize(h.stream));
    return LSS_OK;
}

lss_status lss_autotune(void* handle, const int16_t* d_window, int64_t W, int32_t reps, float* grid_ms,
                        int32_t* best_grid, int32_t* first_passes) {
    if (!handle || !d_window || reps < 1 || !best_grid || !first_passes) return fail(LSS_ERR_ARG, "bad argument");
    Handle& h = *as_handle(handle);
    static const int kGrids[] = {256, 512, 1024, 2048, 4096, 8192};
    lss_spikes out{};
    out.capacity = h.prm.max_spikes;   // counts only: every output array is NULL
    float best = 1e30f;
    int bg = h.upd_grid, bc = h.upd_cpw, max_passes_seen = 0;
    const int pers = h.pers_grid;   // the sweep times the per-pass kernels
    h.pers_grid = 0;
    for (int li = 0; li < 2; ++li) {
        h.upd_cpw = 2 - li;          // two, then one score column per wavefront
        for (int gi = 0; gi < 6; ++gi) {
            h.upd_grid = kGrids[gi];
            float t = 1e30f;
            for (int r = 0; r <= reps; ++r) {   // first run warms up
                if (lss_status st = lss_sort_window(handle, d_window, W, 1, 0, &out)) {
                    h.pers_grid = pers;
                    return st;
                }
                if (r > 0) t = std::min(t, h.last_ms[7]);
                max_passes_seen = std::max(max_passes_seen, out.passes);
            }
            if (grid_ms) grid_ms[6 * li + gi] = t;
            if (t < best) {
                best = t;
                bg = kGrids[gi];
                bc = h.upd_cpw;
            }
        }
    }
    h.pers_grid = pers;
    h.upd_grid = bg;
    h.upd_cpw = bc;
    // the first graph segment covers the passes seen plus a margin; the rest run only when needed
    h.tuned_passes = std::min(h.prm.max_passes, max_passes_seen + 4);
    *best_grid = bg;
    *first_passes = h.tuned_passes;
    return LSS_OK;
}

lss_status lss_set_update_columns(void* handle, int32_t per_wave) {
    if (!handle || (per_wave != 1 && per_wave != 2)) return fail(LSS_ERR_ARG, "bad argument");
    as_handle(handle)->upd_cpw = per_wave;
    return LSS_OK;
}

lss_status lss_get_update_columns(void* handle, int32_t* per_wave) {
    if (!handle || !per_wave) return fail(LSS_ERR_ARG, "null argument");
    *per_wave = as_handle(handle)->upd_cpw;
    return LSS_OK;
}

}  // extern "C"
