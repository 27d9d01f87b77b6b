// Drives the C++ facade (include/lss_online_spikes.hpp) over a recorded stream, as
// OnlineSpikesV2::runSpikeSorting runs over SpikeGLX: a DataSocket serves an int16 [n][C] file on a
// simulated clock that advances only when the sorter waits, so the batches are reproducible.
//
// usage: online_facade_check <oss_input dir> <recording.i16> <n_chan> <min_window> <max_window>
//                            <out dir> [time_behind_ms first_count lag small_skip]
// The clock starts at first_count; each wait moves it to the awaited count and then lag samples
// further (a stream that runs ahead of the sorter).  Writes <out dir>/spikeOutput.txt through the
// facade and prints one line "batches B skips S spikes N processed P channels M first F last L"
// (M channels handed to the socket, F and L the first and last of them).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "lss_online_spikes.hpp"

namespace {

class FileSocket : public lss::DataSocket {
public:
    FileSocket(std::vector<int16_t> data, long C, long first, long lag) : d_(std::move(data)), C_(C), lag_(lag) {
        n_ = (long)(d_.size() / C_);
        now_ = std::min(first, n_);
    }
    lss::t_ull getStreamSampleCt(int, lss::OSSSpecificParams) override { return (lss::t_ull)now_; }
    lss::t_ull fetchFromPlace(int16_t* out, lss::OSSSpecificParams os, lss::t_ull start, long maxs) override {
        chans_ = os.vImecChannels;
        const long s = (long)start;
        const long k = std::max(0L, std::min(maxs, now_ - s));
        std::copy(d_.begin() + s * C_, d_.begin() + (s + k) * C_, out);
        return (lss::t_ull)(s + k);
    }
    void waitUntil(lss::t_ull ct, lss::OSSSpecificParams) override {
        if ((long)ct > n_) ended_ = true;
        now_ = std::max(now_, std::min((long)ct, n_));
        now_ = std::min(now_ + lag_, n_);
    }
    bool isRunning() override { return !ended_; }
    const std::vector<int>& channels() const { return chans_; }

private:
    std::vector<int> chans_;   // the channel list of the last fetch
    std::vector<int16_t> d_;
    long C_, lag_, n_, now_;
    bool ended_ = false;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s dir rec.i16 n_chan min_window max_window out_dir [time_behind_ms first_count lag small_skip]\n",
                     argv[0]);
        return 2;
    }
    const long C = std::atol(argv[3]);
    std::ifstream f(argv[2], std::ios::binary | std::ios::ate);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", argv[2]);
        return 2;
    }
    const long bytes = (long)f.tellg();
    f.seekg(0);
    std::vector<int16_t> rec(bytes / 2);
    f.read((char*)rec.data(), bytes);
    lss::InputParameters p;
    p.sInputFolder = argv[1];
    p.iMinScanWindow = std::atoi(argv[4]);
    p.iMaxScanWindow = std::atoi(argv[5]);
    p.sOSSOutputFolder = argv[6];
    p.iMaxPasses = 100;
    long first = 0, lag = 0;
    if (argc >= 11) {
        p.iTimeBehind = std::atoi(argv[7]);
        first = std::atol(argv[8]);
        lag = std::atol(argv[9]);
        p.bSmallskip = std::atoi(argv[10]) != 0;
    }
    FileSocket sock(std::move(rec), C, first, lag);
    long spikes = 0;
    try {
        lss::OnlineSpikesV2 oss(p, [&](const lss::OnlineSpikesPayload& pl) { spikes += (long)pl.Times.size(); },
                                &sock);
        oss.runSpikeSorting();
        const std::vector<int>& ch = sock.channels();
        std::printf("batches %ld skips %ld spikes %ld processed %ld channels %zu first %d last %d\n", oss.batches(),
                    oss.skips(), spikes, oss.processedCt(), ch.size(), ch.empty() ? -1 : ch.front(),
                    ch.empty() ? -1 : ch.back());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
