// Output step of the sorter, host side: nearest-centroid labels, duplicate removal, text writer.
// Follows OnlineSpikesV2::closestCluster / saveSpikes / writeSpikesToFile
// (src/SpikeSorter/OnlineSpikesV2.cu:1028-1043, 1117-1164) with Kilosort's duplicate rule
// (kilosort/postprocessing.py:10-30: a spike is kept when it is at least dt after the last kept
// spike of its cluster), applied to time-ordered spikes with per-label state kept across windows.
#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/lss.h"

namespace lss {
lss_status set_error(lss_status st, const char* msg);
}

extern "C" {

lss_status lss_label_spikes(const float* x, const float* y, int64_t n, const float* cx, const float* cy,
                            int32_t n_clusters, int32_t* label) {
    if (n < 0 || (n > 0 && (!x || !y || !label)) || n_clusters < 1 || !cx || !cy)
        return lss::set_error(LSS_ERR_ARG, "lss_label_spikes: bad argument");
    for (int64_t i = 0; i < n; ++i) {
        float best = INFINITY;
        int32_t bi = -1;
        for (int32_t t = 0; t < n_clusters; ++t) {
            const float dx = x[i] - cx[t], dy = y[i] - cy[t];
            const float d = std::sqrt(dx * dx + dy * dy);
            if (d < best) {
                best = d;
                bi = t;
            }
        }
        label[i] = bi;
    }
    return LSS_OK;
}

lss_status lss_dedup_spikes(int64_t* sample, int32_t* label, float* amplitude, int64_t n, int32_t dt,
                            int64_t* last_kept, int32_t n_labels, int64_t* n_out) {
    if (n < 0 || (n > 0 && (!sample || !label)) || !last_kept || n_labels < 1 || !n_out || dt < 0)
        return lss::set_error(LSS_ERR_ARG, "lss_dedup_spikes: bad argument");
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (i > 0 && sample[i] < sample[i - 1])
            return lss::set_error(LSS_ERR_ARG, "lss_dedup_spikes: samples not in ascending order");
        const int32_t c = label[i];
        if (c < 0 || c >= n_labels) return lss::set_error(LSS_ERR_ARG, "lss_dedup_spikes: label out of range");
        if (sample[i] < last_kept[c] + dt) continue;
        last_kept[c] = sample[i];
        sample[k] = sample[i];
        label[k] = c;
        if (amplitude) amplitude[k] = amplitude[i];
        ++k;
    }
    *n_out = k;
    return LSS_OK;
}

lss_status lss_write_spikes_txt(const char* path, const int64_t* sample, const int32_t* label,
                                const float* amplitude, const float* y, int64_t n,
                                int64_t recording_offset, int32_t append) {
    if (!path || n < 0 || (n > 0 && (!sample || !label || !amplitude || !y)))
        return lss::set_error(LSS_ERR_ARG, "lss_write_spikes_txt: bad argument");
    FILE* f = std::fopen(path, append ? "a" : "w");
    if (!f) return lss::set_error(LSS_ERR_IO, (std::string("cannot open ") + path).c_str());
    for (int64_t i = 0; i < n; ++i)
        std::fprintf(f, "%lld,%d,%.9g,%.9g\n", (long long)(recording_offset + sample[i]), label[i],
                     (double)amplitude[i], (double)y[i]);
    const bool ok = std::fclose(f) == 0;
    return ok ? LSS_OK : lss::set_error(LSS_ERR_IO, "write failed");
}

}  // extern "C"
