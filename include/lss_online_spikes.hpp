/*
 * lss_online_spikes.hpp - C++ facade with the shape of LiveSpikeSorter's spike-sorter module, over
 * the C ABI of lss.h (header only: include it and link liblss.so).
 *
 *   class OnlineSpikesV2        src/SpikeSorter/OnlineSpikesV2.h:29-34: built from the input
 *                               parameters, a result sink (in place of the decoder's socket address)
 *                               and the shared DataSocket; runSpikeSorting() is the batch loop of
 *                               src/SpikeSorter/OnlineSpikesV2.cu:632-820.
 *   class DataSocket            src/SpikeSorter/dataSocket.h:23-57, the sample source.  It delivers
 *                               int16 samples, interleaved [n][C]: the reference converts the same
 *                               integers to float on the host (dataSocket.cpp:120-122); the library
 *                               converts on the device.
 *   struct OnlineSpikesPayload  src/Networking/onlineSpikesPayload.h:8-49, what one batch hands on.
 *   struct InputParameters      the fields of src/Networking/inputParameters.h:10-74 the sorter
 *                               reads, plus the build's own settings.
 *
 * Batching.  Each batch reports the spikes of n new samples [start, start + n) (minWindow <= n <=
 * maxWindow) from the window [start - nt, start + n + nt): Kilosort's padded batch
 * (kilosort/io.py:584-637), so consecutive batches tile the stream without overlap and no spike is
 * reported twice.  The batch waits until the right-hand halo has arrived.  A sorter that lags the
 * stream follows the reference's policy (OnlineSpikesV2.cu:675-709): timeBehind == 0 jumps to the
 * newest maxWindow samples; within timeBehind ms it catches up with maxWindow-sample batches; past
 * it, smallSkip resumes at the allowed lag and otherwise at the newest batch.  Samples before the
 * stream's first one repeat it (Kilosort's first-batch padding).  Spike times are recording
 * samples of the trough (kilosort/template_matching.py:68), labels the nearest cluster centroid
 * (OnlineSpikesV2.cu:1028-1043; template ids when no centroids are given), and same-label spikes
 * closer than dt samples are dropped with the state carried across batches
 * (kilosort/postprocessing.py:10-30).
 *
 * Errors: as in the reference module (Utils.h:41-64, OnlineSpikesV2.cu:57-65), failures throw
 * std::runtime_error, with the library's message; the C ABI underneath never throws.
 */
#ifndef LSS_ONLINE_SPIKES_HPP_
#define LSS_ONLINE_SPIKES_HPP_

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdint>
#include <functional>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "lss.h"

namespace lss {

typedef unsigned long long t_ull;

enum streamTypes { NIDQ, OBX, IMEC };

struct OSSSpecificParams {
    long lNChans;
    std::vector<int> vImecChannels;
    int substream;
};

class DataSocket {
public:
    virtual ~DataSocket() {}
    /* Stream sample count: samples [0, count) have been acquired. */
    virtual t_ull getStreamSampleCt(int streamType, OSSSpecificParams osParams) = 0;
    /* Samples [lStartCt, lStartCt + k), k = min(lMaxSamples, count - lStartCt), into data as int16
     * [k][lNChans]; returns lStartCt + k. */
    virtual t_ull fetchFromPlace(int16_t* data, OSSSpecificParams osParams, t_ull lStartCt, long lMaxSamples) = 0;
    /* Blocks until the stream count reaches lCt (or the stream ends). */
    virtual void waitUntil(t_ull lCt, OSSSpecificParams osParams) { (void)lCt; (void)osParams; }
    virtual bool isRunning() { return true; }
};

struct InputParameters {
    std::string sInputFolder;            // oss_input directory (lss_create_from_dir)
    std::string sOSSOutputFolder;        // spikeOutput.txt is written here; "" = no file
    unsigned short uSelectedDevice = 0;  // HIP device of this sorter
    int iSubstream = 0;
    int iMinScanWindow = 3000;           // new samples per batch, at least
    int iMaxScanWindow = 3000;           // new samples per batch, at most
    int iTimeBehind = 100000;            // ms the sorter may lag the stream (0: always the newest)
    bool bSmallskip = false;             // past the allowed lag: resume there (true) or at the newest
    float fImecSamplingRate = 30000.f;
    std::vector<int> vChannelSubset;     // channels of this sorter (OSSSpecificParams::vImecChannels)
    // settings of this build
    int iMaxPasses = 50;                 // pursuit pass cap (OnlineSpikesV2.cu:903)
    int iPrecision = LSS_PREC_FP32;
    int iMaxSpikes = 1 << 16;            // spikes one batch may produce
    int iDt = -1;                        // duplicate window in samples; < 0: misc.txt's
                                         // duplicate_spike_bins (OnlineSpikesV2.cu:456)
    std::vector<float> vCentroidX, vCentroidY;   // cluster centroids (um); empty: those of
                                                 // sInputFolder/cluster_centroids.npy, or template ids
};

struct OnlineSpikesPayload {
    long recordingOffset;
    long streamSampleCt;
    std::vector<long> Times;
    std::vector<long> Templates;
    std::vector<float> Amplitudes;
    double VRMS;
    float P2P;
    long processTime;                    // microseconds
};

class OnlineSpikesV2 {
public:
    typedef std::function<void(const OnlineSpikesPayload&)> Sink;

    OnlineSpikesV2(const InputParameters& params, Sink sink, DataSocket* socket)
        : prm_(params), sink_(std::move(sink)), sock_(socket) {
        if (!sock_) throw std::runtime_error("OnlineSpikesV2: null DataSocket");
        if (prm_.iMinScanWindow < 1 || prm_.iMaxScanWindow < prm_.iMinScanWindow)
            throw std::runtime_error("OnlineSpikesV2: need 1 <= iMinScanWindow <= iMaxScanWindow");
        call(lss_select_device(prm_.uSelectedDevice));
        lss_params p;
        call(lss_default_params(&p));
        p.max_passes = prm_.iMaxPasses;
        p.precision = prm_.iPrecision;
        p.max_spikes = prm_.iMaxSpikes;
        p.fs = prm_.fImecSamplingRate;
        p.max_window = prm_.iMaxScanWindow + 2 * 127;   // nt <= 127 comes from the artifacts
        call(lss_create_from_dir(prm_.sInputFolder.c_str(), &p, &h_));
        try {
            setup();
        } catch (...) {
            lss_destroy(h_);
            h_ = nullptr;
            throw;
        }
    }

    ~OnlineSpikesV2() {
        if (h_) lss_destroy(h_);
    }

    OnlineSpikesV2(const OnlineSpikesV2&) = delete;
    OnlineSpikesV2& operator=(const OnlineSpikesV2&) = delete;

    /* Batches until stop() or the end of the stream. */
    void runSpikeSorting() {
        start_clock();
        while (!stop_.load() && processNext() >= 0) {
        }
    }

private:
    void setup() {
        lss_params got;
        call(lss_get_params(h_, &got));
        C_ = got.n_chan;
        nt_ = got.nt;
        T_ = got.n_templates;
        if (prm_.iDt < 0) prm_.iDt = got.dup_bins;
        os_.lNChans = C_;
        if (prm_.vChannelSubset.empty() && !prm_.sInputFolder.empty()) {
            // the artifact directory's channel list (channelMap.npy), when there is one
            int32_t n = 0;
            if (lss_load_channel_map(prm_.sInputFolder.c_str(), nullptr, 0, &n) == LSS_OK && n > 0) {
                if (n != C_)   // OnlineSpikesV2::loadChannelMap requires one entry per channel
                    throw std::runtime_error("OnlineSpikesV2: channelMap.npy has " + std::to_string(n) +
                                             " entries, the templates " + std::to_string(C_) + " channels");
                std::vector<int32_t> m(n);
                call(lss_load_channel_map(prm_.sInputFolder.c_str(), m.data(), n, &n));
                prm_.vChannelSubset.assign(m.begin(), m.end());
            }
        }
        os_.vImecChannels = prm_.vChannelSubset;
        os_.substream = prm_.iSubstream;
        if (prm_.vCentroidX.empty() && !prm_.sInputFolder.empty()) {
            // the artifact directory's label centroids (cluster_centroids.npy, as
            // OnlineSpikesV2::loadKilosortClusteringData reads them); template ids without it
            int32_t n = 0;
            if (lss_load_cluster_centroids(prm_.sInputFolder.c_str(), nullptr, nullptr, 0, &n) == LSS_OK && n > 0) {
                prm_.vCentroidX.resize(n);
                prm_.vCentroidY.resize(n);
                call(lss_load_cluster_centroids(prm_.sInputFolder.c_str(), prm_.vCentroidX.data(),
                                                prm_.vCentroidY.data(), n, &n));
            }
        }
        if (prm_.vCentroidX.size() != prm_.vCentroidY.size())
            throw std::runtime_error("OnlineSpikesV2: centroid x and y differ in length");
        nlabels_ = prm_.vCentroidX.empty() ? T_ : (int)prm_.vCentroidX.size();
        last_kept_.assign(nlabels_, LLONG_MIN / 2);
        const size_t cap = (size_t)prm_.iMaxSpikes;
        sample_.resize(cap);
        tmpl_.resize(cap);
        amp_.resize(cap);
        x_.resize(cap);
        y_.resize(cap);
        // the window buffer is page-locked, as the reference's fetch buffer is (OnlineSpikesV2.cu:399-405):
        // lss_sort_window's upload is then a DMA from it
        void* wp = nullptr;
        call(lss_host_alloc(sizeof(int16_t) * (size_t)(prm_.iMaxScanWindow + 2 * nt_) * C_, &wp));
        win_.reset((int16_t*)wp);
        if (!prm_.sOSSOutputFolder.empty()) {
            out_path_ = prm_.sOSSOutputFolder;
            if (out_path_.back() != '/') out_path_ += '/';
            out_path_ += "spikeOutput.txt";
            call(lss_write_spikes_txt(out_path_.c_str(), nullptr, nullptr, nullptr, nullptr, 0, 0, 0));
        }
    }

public:
    /* One batch: waits for its samples, sorts and hands the spikes on.  Returns the number of
     * spikes handed on, or -1 when the stream has ended before the batch was complete. */
    long processNext() {
        start_clock();
        const long nt = nt_, mn = prm_.iMinScanWindow, mx = prm_.iMaxScanWindow;
        sock_->waitUntil((t_ull)(processed_ + mn + nt), os_);
        const auto t0 = std::chrono::steady_clock::now();
        const long latest = (long)sock_->getStreamSampleCt(IMEC, os_);
        if (latest < processed_ + mn + nt) return -1;
        const long ready = latest - nt;                       // last reportable sample + 1
        const long allowed = ready - (long)((double)prm_.iTimeBehind * prm_.fImecSamplingRate / 1000.0);
        long start = processed_, n = ready - processed_;
        bool skipped = false;
        if (n > mx) {
            if (prm_.iTimeBehind == 0) {
                start = ready - mx;
                skipped = true;
            } else if (processed_ >= allowed) {
                // lagging within the tolerance: catch up from here
            } else if (prm_.bSmallskip) {
                start = std::min(allowed, ready - mx);
                skipped = true;
            } else {
                start = ready - mx;
                skipped = true;
            }
            n = mx;
        }
        if (skipped) ++skips_;
        const long W = n + 2 * nt;
        fetch_window(start - nt, W);
        float p2p = 0.f;
        {
            int16_t lo = INT16_MAX, hi = INT16_MIN;
            const int16_t* q = win_.get() + (size_t)nt * C_;
            for (size_t i = 0; i < (size_t)n * C_; ++i) {
                lo = std::min(lo, q[i]);
                hi = std::max(hi, q[i]);
            }
            p2p = (float)hi - (float)lo;
        }
        lss_spikes out{};
        out.capacity = prm_.iMaxSpikes;
        out.sample = sample_.data();
        out.template_id = tmpl_.data();
        out.amplitude = amp_.data();
        out.x = x_.data();
        out.y = y_.data();
        call(lss_sort_window(h_, win_.get(), W, 0, start - nt, &out));
        const int64_t ns = out.n;
        // labels, time order, duplicates
        std::vector<int32_t> lab((size_t)ns);
        if (prm_.vCentroidX.empty()) {
            std::copy(tmpl_.begin(), tmpl_.begin() + ns, lab.begin());
        } else if (ns > 0) {
            call(lss_label_spikes(x_.data(), y_.data(), ns, prm_.vCentroidX.data(), prm_.vCentroidY.data(),
                                  nlabels_, lab.data()));
        }
        std::vector<int64_t> ord((size_t)ns);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return sample_[a] < sample_[b]; });
        std::vector<int64_t> s((size_t)ns);
        std::vector<int32_t> l((size_t)ns);
        std::vector<float> idx((size_t)ns);     // original positions ride along as the dedup's float column
        for (int64_t i = 0; i < ns; ++i) {
            s[i] = sample_[ord[i]];
            l[i] = lab[ord[i]];
            idx[i] = (float)ord[i];
        }
        int64_t kept = 0;
        call(lss_dedup_spikes(s.data(), l.data(), idx.data(), ns, prm_.iDt, last_kept_.data(), nlabels_, &kept));
        OnlineSpikesPayload pl;
        pl.recordingOffset = recording_offset_;
        pl.streamSampleCt = latest;
        pl.VRMS = 0.0;
        pl.P2P = p2p;
        std::vector<float> a((size_t)kept), yy((size_t)kept);
        for (int64_t i = 0; i < kept; ++i) {
            const int64_t o = (int64_t)idx[i];
            pl.Times.push_back((long)s[i]);
            pl.Templates.push_back((long)l[i]);
            pl.Amplitudes.push_back(amp_[o]);
            a[i] = amp_[o];
            yy[i] = y_[o];
        }
        if (!out_path_.empty())
            call(lss_write_spikes_txt(out_path_.c_str(), s.data(), l.data(), a.data(), yy.data(), kept,
                                      recording_offset_, 1));
        processed_ = start + n;
        ++batches_;
        pl.processTime = (long)std::chrono::duration_cast<std::chrono::microseconds>(
                             std::chrono::steady_clock::now() - t0).count();
        if (sink_) sink_(pl);
        return (long)kept;
    }

    void stop() { stop_.store(true); }
    long batches() const { return batches_; }
    long skips() const { return skips_; }
    /* First sample not yet reported. */
    long processedCt() const { return processed_; }
    void* handle() const { return h_; }

private:
    static void call(lss_status st) {
        if (st != LSS_OK) throw std::runtime_error(std::string("liblss: ") + lss_last_error());
    }

    // the reference starts at the stream's current count (OnlineSpikesV2.cu:667-668)
    void start_clock() {
        if (init_) return;
        processed_ = (long)sock_->getStreamSampleCt(IMEC, os_);
        init_ = true;
    }

    // samples [first, first + W) into win_; rows before sample 0 repeat sample 0
    void fetch_window(long first, long W) {
        int16_t* w = win_.get();
        long row = 0;
        if (first < 0) {
            const long pad = std::min(-first, W);
            sock_->fetchFromPlace(w + (size_t)pad * C_, os_, 0, 1);
            for (long r = 0; r < pad; ++r)
                std::copy(w + (size_t)pad * C_, w + (size_t)(pad + 1) * C_, w + (size_t)r * C_);
            row = pad;
        }
        while (row < W) {
            const t_ull got = sock_->fetchFromPlace(w + (size_t)row * C_, os_, (t_ull)(first + row), W - row);
            const long k = (long)got - (first + row);
            if (k <= 0) throw std::runtime_error("OnlineSpikesV2: DataSocket returned no samples");
            row += k;
        }
    }

    InputParameters prm_;
    Sink sink_;
    DataSocket* sock_;
    OSSSpecificParams os_{};
    void* h_ = nullptr;
    int C_ = 0, nt_ = 0, T_ = 0, nlabels_ = 0;
    long processed_ = 0, batches_ = 0, skips_ = 0, recording_offset_ = 0;
    bool init_ = false;
    std::atomic<bool> stop_{false};
    std::string out_path_;
    std::vector<int64_t> last_kept_;
    std::vector<int64_t> sample_;
    std::vector<int32_t> tmpl_;
    std::vector<float> amp_, x_, y_;
    struct HostFree {
        void operator()(int16_t* p) const { (void)lss_host_free(p); }
    };
    std::unique_ptr<int16_t, HostFree> win_;   // page-locked window [iMaxScanWindow + 2 nt][C]
};

}  // namespace lss

#endif
