/*
 * lss.h - C ABI of the MI355X Kilosort4 online template-matching library (liblss.so).
 *
 * Replaces the per-batch operator of LiveSpikeSorter's SpikeSorter module:
 *   OnlineSpikesV2(InputParameters, sockaddr_in, DataSocket*)      src/SpikeSorter/OnlineSpikesV2.h:32
 *   long kilosortMatchingPursuit(float* d_batch, long W)          src/SpikeSorter/OnlineSpikesV2.h:56
 *   void computeClosestClusters(long W, long nsp)                 src/SpikeSorter/OnlineSpikesV2.h:57
 *   void highpassFilter(float*, int C, int W, float fs, float f)  src/SpikeSorter/OnlineSpikesV2.h:58
 *   meanRemove / medianRemove / projectToPCA / crossCorrelation /
 *   normalizeConv / reduceToTimeDimByMax / findMatchingIndices /
 *   updateResidual                                                 src/SpikeSorter/myGPUhelpers.h:20-49
 * with the semantics of the Kilosort4 Python the module reimplements
 * (kilosort/io.py:584-637, 757-790; kilosort/template_matching.py:20-26, 50-51, 110-121, 128-202;
 * kilosort/postprocessing.py:33-46).
 *
 * Conventions: plain pointers and sizes, no exceptions and no exit() across the ABI.  Every call
 * returns an lss_status; lss_last_error() gives the message of the last failure on the calling
 * thread.  A handle owns all device memory it uses, sized once at creation for max_window; it is
 * not thread-safe and issues all work on its own HIP stream (lss_set_stream to share one).
 * Device pointers passed to stage entry points must be device memory of the handle's GPU.
 */
#ifndef LSS_H_
#define LSS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSS_ABI_VERSION 2

typedef enum lss_status {
    LSS_OK = 0,
    LSS_ERR_ARG = 1,        /* bad argument / shape */
    LSS_ERR_HIP = 2,        /* HIP runtime error */
    LSS_ERR_IO = 3,         /* artifact directory / .npy problem */
    LSS_ERR_CAPACITY = 4,   /* more spikes than the output capacity */
    LSS_ERR_STATE = 5       /* handle not usable */
} lss_status;

/* Arithmetic of the two large contractions (whitening and U.Xproj). */
typedef enum lss_precision {
    LSS_PREC_FP32 = 0,      /* FP32 MFMA, FP32 accumulation */
    LSS_PREC_BF16 = 1,      /* BF16 MFMA, FP32 accumulation */
    LSS_PREC_FP16 = 2,      /* FP16 MFMA, FP32 accumulation */
    LSS_PREC_BF16X3 = 3,    /* split-BF16 (hi*hi + hi*lo + lo*hi) on BF16 MFMA */
    LSS_PREC_FP16X3 = 4     /* split-FP16 (hi*hi + hi*lo + lo*hi) on FP16 MFMA */
} lss_precision;

/* Sorter configuration (OnlineSpikesV2 members T, K, M, C, nt0min, numNearestChans, Th_learned,
 * dt: src/SpikeSorter/OnlineSpikesV2.h:84-99; misc.txt keys OnlineSpikesV2.cu:435-457). */
typedef struct lss_params {
    int32_t n_chan;        /* C */
    int32_t n_templates;   /* T (<= 2048; above 1024 the pursuit always runs on per-pass kernels) */
    int32_t n_pcs;         /* K (<= 8) */
    int32_t nt;            /* template length M (odd, <= 127) */
    int32_t nt0min;        /* trough sample of the template window */
    int32_t n_nearest;     /* nC channels of the PC features (<= 32) */
    int32_t max_window;    /* largest padded window W' (samples) a call may pass */
    int32_t max_passes;    /* pursuit pass cap: 50 (OnlineSpikesV2.cu:903) or 100 (template_matching.py:157) */
    int32_t max_spikes;    /* spikes one window may produce (device list capacity) */
    int32_t do_car;        /* median (common average) reference, kilosort/io.py:766-768 */
    int32_t precision;     /* lss_precision */
    int32_t drop_negative; /* drop spikes whose sample < 0 (template_matching.py:53-59) */
    double th_learned;     /* Th (Kilosort's Th_learned, a double): a peak needs relu(B)^2/nm >
                              (float)(Th*Th), the reference's FP32 comparison with Th**2 */
    float fs;              /* sampling rate (Hz) */
    int32_t dup_bins;      /* duplicate window of the output step in samples (misc.txt
                              duplicate_spike_bins, OnlineSpikesV2.cu:456; default 7) */
} lss_params;

/* Kilosort4 outputs as host float32 row-major arrays (the oss_input files,
 * run_online_spikes.py:545-576). */
typedef struct lss_artifacts {
    const float* U;          /* [T][K][C]  Wall3.npy */
    const float* wPCA;       /* [K][nt]    wPCA.npy */
    const float* Wrot;       /* [C][C]     whiteningMat.npy */
    const float* drift;      /* [C][C]     drift matrix M applied as (M @ Wrot), or NULL for I */
    const float* hp_filter;  /* [hp_len]   hp_filter.npy (filtfilt impulse response) */
    int32_t hp_len;
    const float* xc;         /* [C] channel x (um) */
    const float* yc;         /* [C] channel y (um) */
    /* Optional feature-extraction tables of prepare_extract (kilosort/template_matching.py:20-26),
     * as the exporter writes them (run_online_spikes.py:533, 572-574) and OnlineSpikesV2 loads them
     * (OnlineSpikesV2.cu:561-563); NULL = built from U and the channel sites (iCC with ties broken
     * by channel index).  A template refinement step rebuilds iU and Ucc from the refined U. */
    const int32_t* iCC;      /* [nC][C] nearest channels of each channel, or NULL */
    const int32_t* iU;       /* [T] channel of largest template energy, or NULL */
    const float* Ucc;        /* [nC][T][K] U at the nearest channels of iU, or NULL */
} lss_artifacts;

/* Caller-owned host output buffers of one window. */
typedef struct lss_spikes {
    int64_t capacity;        /* rows available in every non-NULL array */
    int64_t n;               /* out: spikes written */
    int32_t passes;          /* out: pursuit passes that found spikes */
    int64_t* sample;         /* [capacity] recording sample of the trough */
    int32_t* template_id;    /* [capacity] */
    float* amplitude;        /* [capacity] B[t,w]/nm[t] */
    float* features;         /* [capacity][nC][K] PC features, or NULL */
    float* x;                /* [capacity] position (um), or NULL */
    float* y;                /* [capacity] */
    int32_t* column;         /* [capacity] padded-window column of the peak, or NULL */
} lss_spikes;

const char* lss_last_error(void);
int32_t lss_abi_version(void);

/* Fills defaults: nt=61, nt0min=20, K=6, nC=10, Th=8, fs=30000, 50 passes, FP32, CAR on,
 * drop_negative on, max_window=60122. */
lss_status lss_default_params(lss_params* p);

/* Creates a sorter on the current HIP device: uploads the artifacts and builds the FP64-derived
 * tables (template norms, template waveforms, ctc) on the device
 * (replaces OnlineSpikesV2::initializeSorter / initializeStaticMemory, OnlineSpikesV2.cu:321-630). */
lss_status lss_create(const lss_params* p, const lss_artifacts* a, void** handle);

/* Same, reading an oss_input directory as OnlineSpikesV2::loadKilosortTrainingData does
 * (OnlineSpikesV2.cu:435-457, 546-610): Wall3.npy, wPCA.npy, whiteningMat.npy or whitening_mat.npy,
 * drift_matrix.npy, hp_filter.npy, xc.npy, yc.npy, misc.txt (nt0min, numNearestChans, Th_learned,
 * duplicate_spike_bins), and when present iCC.npy [nC][C], iU.npy [T] (any integer dtype) and
 * Ucc.npy [nC][T][K], which are used as loaded.  A present ctc.npy [T][T][2nt+1] is compared with
 * the table rebuilt in FP64 (lss_get_load_info); the rebuilt one is used.  p supplies everything
 * misc.txt and the array shapes do not. */
lss_status lss_create_from_dir(const char* dir, const lss_params* p, void** handle);

/* What lss_create / lss_create_from_dir took from the caller or the directory. */
typedef struct lss_load_info {
    int32_t icc_given;         /* iCC came from the caller / iCC.npy (else built) */
    int32_t iu_given;          /* iU likewise */
    int32_t ucc_given;         /* Ucc likewise */
    int32_t ctc_in_dir;        /* ctc.npy was present and compared */
    double ctc_max_abs_diff;   /* max |ctc.npy - FP64 rebuild| (0 without ctc.npy) */
    double ctc_max_abs;        /* max |FP64 rebuild| */
} lss_load_info;
lss_status lss_get_load_info(void* handle, lss_load_info* info);

lss_status lss_destroy(void* handle);
/* Makes `device` the calling thread's current HIP device (OnlineSpikesV2::initializeSorter's
 * cudaSetDevice, OnlineSpikesV2.cu:326-327); handles are created on the current device. */
lss_status lss_select_device(int32_t device);
/* Page-locked host memory for windows and sample chunks: lss_sort_window and lss_stream_push read
 * it by DMA at the link's full rate, where pageable memory is staged through the runtime's own
 * buffers.  The reference keeps its fetch buffer page-locked the same way (its "Pinned" memory
 * variables, cudaMallocHost, src/SpikeSorter/OnlineSpikesV2.cu:399-405). */
lss_status lss_host_alloc(size_t bytes, void** p);
lss_status lss_host_free(void* p);
lss_status lss_get_params(void* handle, lss_params* p);
lss_status lss_set_stream(void* handle, void* hip_stream);

/* One padded window int16 [W][C] (interleaved, as fetched) through the whole path:
 * CAR, highpass, whitening, projection, pursuit, features, positions.  window_start is the
 * recording sample of column 0 (ibatch*NT - nt for Kilosort batching), so a spike's sample is
 * window_start + column - nt/2 + nt0min (template_matching.py:68).  window may be host or device
 * memory (on_device).  Replaces kilosortMatchingPursuit + computeClosestClusters. */
lss_status lss_sort_window(void* handle, const int16_t* window, int64_t W, int32_t on_device,
                           int64_t window_start, lss_spikes* out);

/* Several windows with up to n_handles of them in flight: window i is sorted by handles[i %
 * n_handles] (distinct handles of one device, one host thread each, windows of a handle in
 * order), so one window's GEMMs overlap another's pursuit.  outs[i] receives window i's spikes as
 * lss_sort_window gives them.  Returns the first failure (its message names the window). */
lss_status lss_sort_windows(void* const* handles, int32_t n_handles, const int16_t* const* windows, const int64_t* W,
                            const int64_t* window_start, int32_t on_device, int32_t n_windows, lss_spikes* outs);
/* Kilosort batch padding of a host int16 recording [n_samples][C] (kilosort/io.py:584-637):
 * writes the padded window [NT+2nt][C] and its window_start.  Host only. */
lss_status lss_pad_batch(const int16_t* rec, int64_t n_samples, int32_t C, int32_t NT, int32_t nt,
                         int64_t ibatch, int16_t* out, int64_t* window_start);

/* ---- stage entry points (device pointers, handle's stream, synchronous on return) ---------- */

/* int16 window [W][C] (device) -> X after mean, median, highpass: [C][W] (before whitening). */
lss_status lss_stage_filter(void* handle, const int16_t* d_window, int64_t W, float* d_out_cw);
/* int16 window [W][C] (device) -> after mean and median removal only: [C][W]. */
lss_status lss_stage_car(void* handle, const int16_t* d_window, int64_t W, float* d_out_cw);
/* [C][W] -> highpassed [C][W] (in and out may not alias). */
lss_status lss_stage_highpass(void* handle, const float* d_in_cw, int64_t W, float* d_out_cw);
/* [C][W] -> whitened (M @ Wrot) X, time-major [W][C]. */
lss_status lss_stage_whiten(void* handle, const float* d_in_cw, int64_t W, float* d_out_wc);
/* whitened X [W][C] -> scores B [W][T] = U . (wPCA-projected X). */
lss_status lss_stage_scores(void* handle, const float* d_x_wc, int64_t W, float* d_b_wt);
/* Pursuit on given scores B [W][T] and residual X [W][C] (both updated in place), then features
 * and positions from the final residual; spikes in pursuit order (pass, then column). */
lss_status lss_stage_pursuit(void* handle, float* d_b_wt, float* d_x_wc, int64_t W,
                             int64_t window_start, lss_spikes* out);

/* Copies the device tables to host: nm [T], waves [T][nt][C], ctc [T][2nt+1][T]
 * (ctc[j][l][i] = Kilosort ctc[i][j][l]), wmat [C][C] = M @ Wrot; any may be NULL. */
lss_status lss_get_tables(void* handle, float* nm, float* waves, float* ctc, float* wmat);

/* Times of the last lss_sort_window's stages in ms (HIP events on the handle's stream):
 * [0] ingest + CAR, [1] highpass, [2] whitening GEMM, [3] wPCA projection, [4] U.Xproj GEMM,
 * [5] pursuit, [6] features + positions, [7] total; n <= 8. */
lss_status lss_last_timings(void* handle, float* ms, int32_t n);

/* ---- output step (host only; OnlineSpikesV2::closestCluster / saveSpikes / writeSpikesToFile,
 * src/SpikeSorter/OnlineSpikesV2.cu:1028-1043, 1117-1164; Kilosort remove_duplicates,
 * kilosort/postprocessing.py:10-30) ------------------------------------------------------------- */

/* label[i] = index of the cluster centroid (cx, cy) nearest to spike position (x[i], y[i])
 * (Euclidean, first index on ties). */
lss_status lss_label_spikes(const float* x, const float* y, int64_t n, const float* cx, const float* cy,
                            int32_t n_clusters, int32_t* label);

/* In-place removal of same-label spikes closer than dt samples to the previous kept spike of that
 * label.  Spikes must be in ascending sample order.  last_kept [n_labels] carries the last kept
 * sample of each label across calls (initialise to INT64_MIN/2 before the first window);
 * amplitude may be NULL.  *n_out = spikes kept. */
lss_status lss_dedup_spikes(int64_t* sample, int32_t* label, float* amplitude, int64_t n, int32_t dt,
                            int64_t* last_kept, int32_t n_labels, int64_t* n_out);

/* Appends "sample,label,amplitude,y" lines (sample + recording_offset) to a text file, as the
 * reference's spikeOutput.txt.  append = 0 truncates first. */
lss_status lss_write_spikes_txt(const char* path, const int64_t* sample, const int32_t* label,
                                const float* amplitude, const float* y, int64_t n,
                                int64_t recording_offset, int32_t append);

/* The label centroids of an oss_input directory, cluster_centroids.npy stored flattened as
 * [x1, y1, x2, y2, ...] (OnlineSpikesV2::loadKilosortClusteringData, OnlineSpikesV2.cu:612-628;
 * any numeric dtype, C or Fortran order).  *n = the number of centroids; with x or y NULL only
 * *n is set.  LSS_ERR_IO when the file is missing or malformed, LSS_ERR_CAPACITY when capacity
 * is below *n.  Host only. */
lss_status lss_load_cluster_centroids(const char* dir, float* x, float* y, int32_t capacity, int32_t* n);

/* The acquisition channels of an oss_input directory, channelMap.npy (the channel list
 * OnlineSpikesV2 hands its DataSocket, OnlineSpikesV2.cu:474, 657-661; any integer or float
 * dtype holding non-negative integers).  *n = the number of entries; with out NULL only *n is set.
 * LSS_ERR_IO when the file is missing or malformed, LSS_ERR_CAPACITY when capacity is below *n.
 * Host only. */
lss_status lss_load_channel_map(const char* dir, int32_t* out, int32_t capacity, int32_t* n);

/* ---- feature-space cluster assignment of the reference's Python online sorter
 * (src/Python/OnlineKilosort.py:211-284); device pointers, synchronous ---------------------------- */

/* d_out [T][n_nearest * n_pcs]: each template's PC features on its n_nearest channels (row t =
 * Ucc[:, t, :] flattened channel-major), the cluster centroids of cluster_centroids_pca_compute
 * (OnlineKilosort.py:211-231) with pc_feature_ind[t] = the template's nearest channels. */
lss_status lss_cluster_centroids_pca(void* handle, float* d_out);

/* assign_to_clusters (OnlineKilosort.py:233-284): d_label[i] = the cluster k maximising
 * 2 <tF[i], mu[k]> - |mu[k]|^2 among those with (x[i] - cx[k])^2 + (y[i] - cy[k])^2 <= R^2 (NaN
 * centroids never qualify), the first such k on ties, -1 when none is in range.  tF [n][D]
 * (a spike's [n_nearest][n_pcs] features flattened), mu [n_clusters][D], D <= 256; runs on
 * `stream` (a hipStream_t, NULL = the null stream) and waits for it. */
lss_status lss_assign_clusters(const float* d_tF, int64_t n, int32_t D, const float* d_mu, int32_t n_clusters,
                               const float* d_x, const float* d_y, const float* d_cx, const float* d_cy, double R,
                               int32_t* d_label, void* stream);

/* ---- multi-probe streaming (OnlineSpikesV2::runSpikeSorting's batching loop, V2:632-820, for
 * several probes on one sample clock) ------------------------------------------------------------ */

/* Creates a batcher over n_probes sorter handles (one per probe, distinct).  Each step sorts the
 * next `batch` samples of every probe: probe p's window is samples [t0 - nt, t0 + batch + nt)
 * (samples before 0 repeat sample 0), so consecutive steps report disjoint sample ranges.  With
 * use_graph the device work of a step is replayed HIP graphs whose branches run the probes
 * concurrently.  When every probe runs the persistent pursuit (lss_set_persistent, windows up to
 * LSS_PERSIST_MAX_WINDOW) a step is ONE graph: upload, preprocessing, scores, all pursuit passes,
 * residual, features and a readback of only the spikes found; the probes share the device's
 * co-resident workgroups equally.  Otherwise a step is two graphs, (upload, preprocessing,
 * scores, first P1 passes) and (residual, features, readback), and probes that have not converged
 * after P1 passes (lss_autotune / lss_set_tuning) get their remaining passes in between.  Each
 * handle needs max_window >= batch + 2*nt. */
lss_status lss_stream_create(void* const* handles, int32_t n_probes, int32_t batch, int32_t use_graph,
                             void** stream);
lss_status lss_stream_destroy(void* stream);
/* Appends n samples (int16 [n][C] interleaved, host memory) of one probe. */
lss_status lss_stream_push(void* stream, int32_t probe, const int16_t* samples, int64_t n);
/* Appends n samples of one probe that are already in device memory (int16 [n][C], this GPU): the
 * step copies its window rows device to device and its graph has no upload.  The caller keeps the
 * buffer valid until every step that reads its samples has returned.  A probe takes either host
 * or device samples. */
lss_status lss_stream_push_device(void* stream, int32_t probe, const int16_t* d_samples, int64_t n);
/* The batcher's streams are blocking streams: a step's copies of device samples follow work
 * submitted earlier to the legacy default stream; samples written on any other stream must be
 * complete (that stream synchronised) before the step that reads them. */
/* Runs one batch when every probe holds enough samples (*ready = 1) and fills outs[0..n_probes);
 * otherwise *ready = 0 and nothing runs. */
lss_status lss_stream_step(void* stream, lss_spikes* outs, int32_t* ready);
/* Wall time of the last step in ms (upload to spike lists) and steps done. */
lss_status lss_stream_stats(void* stream, float* last_ms, int64_t* batches);
/* Host wall-time split of the last step in ms: [0] window assembly, [1] segment A (launch to
 * completion), [2] extra passes enqueue, [3] segment C (launch to completion), [4] spike emission,
 * [5] total. */
lss_status lss_stream_breakdown(void* stream, float* ms);
/* Template refinement while streaming: every `every_batches` steps (0 = never) a refinement round
 * starts (unless one is still under way, which carries on) in which probe i's handle takes one
 * lss_refine_templates step on the window of the round's i-th step (drift tracking), so every
 * probe is refined once per round.  The step copies that window's residual and spikes and the
 * refinement runs in the background (a host thread and a stream of the handle's own, writing a
 * second template set); the fourth step after it adopts it at its start, waiting for it if needed,
 * so results do not depend on timing (the steps in between still use the old templates).
 * lss_stream_last_refine gives a probe's last info (4 doubles; it waits for a refinement still
 * under way). */
lss_status lss_stream_set_refine(void* stream, int32_t every_batches);
lss_status lss_stream_last_refine(void* stream, int32_t probe, double* info);
/* Running ahead (on by default; one-graph steps only): a step whose next batch's samples are
 * already pushed also preprocesses that batch (window, whitening, scores) into a second window set
 * per handle, as a second graph beside its own, so that the next step starts with the pursuit; a
 * backlog is then sorted faster.  Results are the same either way.  A step
 * that ends with a refinement does not run ahead, and a batch preprocessed ahead is preprocessed
 * again when lss_set_drift or lss_refine_templates changed its handle in between.  on = 0 turns it
 * off (every step preprocesses its own batch; the per-batch latency of a live stream).  No
 * counterpart in the reference (its loop sorts one batch at a time, V2:632-820). */
lss_status lss_stream_set_run_ahead(void* stream, int32_t on);
/* Steps so far that preprocessed their next batch (running ahead). */
lss_status lss_stream_ran_ahead(void* stream, int64_t* steps);
/* Grouped steps (on by default): when every probe's pursuit runs in the persistent kernel from its
 * first pass (one-graph steps, at most 8 probes, all the probes' persistent workgroups resident at
 * once), a step runs the pursuit, residual, features and readback of all probes as one launch per
 * stage on one stream instead of a branch per probe, which cuts the host cost of launching a step.
 * Results are the same either way.  Set before the first step (LSS_ERR_STATE after it);
 * lss_stream_grouped reports whether the steps run grouped (decided at the first step).  No
 * counterpart in the reference (its loop sorts one probe per process, V2:632-820). */
lss_status lss_stream_set_grouped(void* stream, int32_t on);
lss_status lss_stream_grouped(void* stream, int32_t* grouped);
/* Measurement mode of grouped steps (off by default): each step launches its pursuit tail without
 * the graph, the grouped persistent pursuit's workgroups stamp their GPU-clock start and end, and
 * lss_stream_last_profile gives, for the last step, that launch's span in microseconds (last
 * workgroup end - first workgroup start) and per probe (arrays of n_probes) the spikes found, the
 * passes run and the workgroups the probe took.  Results are the same.  LSS_ERR_STATE when
 * measurement is off or the steps do not run grouped. */
lss_status lss_stream_profile(void* stream, int32_t on);
lss_status lss_stream_last_profile(void* stream, double* gpu_us, int32_t* spikes, int32_t* passes, int32_t* grid);
/* Pipelined steps (on by default; grouped steps running ahead, not in measurement mode): a step that
 * preprocessed its next batch also launches that batch's pursuit before it returns, so the device
 * works on it while the caller handles this step's spikes; the next step then only waits for it.
 * The next step sorts its batch again instead (same results) when, in between, any entry point took
 * one of the stream's handles (each such call first waits for the launch) or an input of its
 * preprocessing changed; a step that starts a refinement or precedes an adoption launches nothing
 * early.  lss_stream_pipelined counts the early launches.  No counterpart in the reference. */
lss_status lss_stream_set_pipelined(void* stream, int32_t on);
lss_status lss_stream_pipelined(void* stream, int64_t* steps);

/* ---- template refinement (no reference counterpart) ---------------------------------------------- */

/* Residual energy of a fixed spike set, the loss of template refinement against drift:
 *   R = X - sum_s amp[s] * wave_{tmp[s]} placed at columns col[s] - nt/2 .. col[s] + nt/2,
 *   wave_t[m][c] = sum_k U[t][k][c] wPCA[k][m]  (the residual update of template_matching.py:197;
 *   contributions of overlapping spikes add; columns outside [0, W) are dropped),
 *   loss = sum R^2 (FP64 sum of FP32 residuals).
 * d_x: [W][C] whitened window (device); d_col/d_tmp/d_amp: n spikes sorted by column (device;
 * col = peak column in the window, as lss_spikes.column); d_U [T][K][C] and d_wpca [K][nt]
 * (device) may differ from the handle's templates.  The residual is recomputed on every call
 * (nothing of a forward pass is kept).  Results are deterministic.  Synchronises the stream. */
lss_status lss_residual_energy(void* handle, const float* d_x, int64_t W, const int32_t* d_col,
                               const int32_t* d_tmp, const float* d_amp, int64_t n, const float* d_U,
                               const float* d_wpca, double* loss);
/* The same loss and its gradient (adjoint): d_grad_U [T][K][C] = dL/dU and d_grad_wpca [K][nt] =
 * dL/dwPCA (device, FP32), from the spike-aligned residual windows
 * G_t[m][c] = sum_{s: tmp=t} amp_s R[col_s - nt/2 + m][c]:
 * dL/dU[t] = -2 wPCA G_t, dL/dwPCA = -2 sum_t U_t G_t. */
lss_status lss_residual_energy_grad(void* handle, const float* d_x, int64_t W, const int32_t* d_col,
                                    const int32_t* d_tmp, const float* d_amp, int64_t n, const float* d_U,
                                    const float* d_wpca, double* loss, float* d_grad_U, float* d_grad_wpca);
/* Device time of the last refinement call in ms: [0] waves + residual + loss, [1] adjoint, [2] total. */
lss_status lss_last_refine_timings(void* handle, float* ms);
/* One refinement step of the handle's own templates on the last window it sorted: the gradient of
 * that window's residual energy (its spikes, the residual the pursuit left) by the adjoint
 * kernel, a step along the negative gradient whose length minimises a quadratic fit of the loss
 * through one trial step, and every derived table (nm, waves, ctc in FP64, iU, Ucc) rebuilt.  The
 * step is adopted only if it lowers the loss.  info (4 doubles, may be NULL): loss before, loss
 * after, step length, gradient norm (zeros when nothing changed).  Synchronises the stream. */
lss_status lss_refine_templates(void* handle, double* info);

/* ---- tuning ------------------------------------------------------------------------------------- */

/* Per-pass kernels: times the window (device memory) with each candidate update layout (2 or 1
 * score columns per wavefront: fewer or lighter work items) and grid (256 .. 8192 workgroups) and
 * keeps the fastest; sets the batcher's first-segment pass count to the passes seen + 4 (capped at
 * max_passes).  grid_ms (12 floats, may be NULL) receives the best time per candidate,
 * [6 * layout + grid] with layout 0 = two columns, 1 = one column per wavefront.  Results do not
 * depend on the tuning. */
lss_status lss_autotune(void* handle, const int16_t* d_window, int64_t W, int32_t reps, float* grid_ms,
                        int32_t* best_grid, int32_t* first_passes);
lss_status lss_set_tuning(void* handle, int32_t update_grid, int32_t first_passes);
lss_status lss_get_tuning(void* handle, int32_t* update_grid, int32_t* first_passes);
/* The update layout lss_autotune chose: score columns per wavefront, 1 or 2 (default 2). */
lss_status lss_set_update_columns(void* handle, int32_t per_wave);
lss_status lss_get_update_columns(void* handle, int32_t* per_wave);

/* Tile shapes of the window's three contractions: the whitening GEMM, the wPCA projection and the
 * U.Xproj score GEMM.  GEMM output tiles: 0 = 128 x 128 (default), 1 = 128 x 64, 2 = 64 x 128,
 * 3 = 64 x 64 samples x outputs (small tiles fill the device on short windows); projection
 * workgroups: FP32 path 0 = by window length (default), 1/2/3 = 32/64/128 samples; split-precision
 * path 0 and 2 = 64 channels (default), 1 = 32, 3 = 128.  Every shape gives bit-identical results
 * (each output's accumulation order is the same), so tuning never changes a spike list.  A
 * streaming batcher captures the shapes set when it first runs a step. */
#define LSS_TILE_SHAPES 4
lss_status lss_set_tiles(void* handle, int32_t whiten, int32_t project, int32_t scores);
/* tiles[3] = whitening, projection, scores */
lss_status lss_get_tiles(void* handle, int32_t* tiles);
/* Times each stage of a real window (device int16 [W][C]) with every shape (HIP events on the
 * handle's stream, best of reps runs after a warm-up run) and keeps the fastest per stage; best[3]
 * receives the choice, ms (12 floats, may be NULL) the times: [4*stage + shape], stage 0 =
 * whitening, 1 = projection, 2 = scores. */
lss_status lss_autotune_tiles(void* handle, const int16_t* d_window, int64_t W, int32_t reps, float* ms,
                              int32_t* best);

/* Persistent pursuit: windows of at most LSS_PERSIST_MAX_WINDOW samples (streaming batches) run
 * their first `first_passes` passes (the ones with many spikes, whose wide updates want the whole
 * chip) as per-pass kernels and the rest in one launch of `grid` co-resident workgroups with two
 * grid barriers per pass, instead of three kernels per pass.  Results are identical.  grid = 0
 * selects the per-pass kernels throughout; grid is clamped at the number of workgroups the device
 * holds at once (max_grid from lss_get_persistent; the streaming batcher gives probe i the
 * fraction sqrt(T_i C_i) / sum_j sqrt(T_j C_j) of its limit, so that all probes fit at once).  A
 * barrier that waits past ~1 s (workgroups not co-resident) makes the call fail with
 * LSS_ERR_STATE instead of hanging.  Defaults: grid 128, first_passes 0. */
#define LSS_PERSIST_MAX_WINDOW 4096
lss_status lss_set_persistent(void* handle, int32_t grid, int32_t first_passes);
/* Current setting and the device's co-resident limit for windows of W samples (0 when W exceeds
 * LSS_PERSIST_MAX_WINDOW). */
lss_status lss_get_persistent(void* handle, int64_t W, int32_t* grid, int32_t* first_passes, int32_t* max_grid);
/* Times the window for grid in {0, 64, 128, 256, 384} x first_passes in {0, 2, 4, P1} (P1 = the
 * first-segment pass count lss_autotune chose, skipped when unset): ms[4*gi + fi]
 * (20 floats, may be NULL; a grid above max_grid or the device limit runs once at that limit;
 * -1 = not run; max_grid 0 = device limit) and keeps the fastest as the handle's setting. */
lss_status lss_autotune_persistent(void* handle, const int16_t* d_window, int64_t W, int32_t reps,
                                   int32_t max_grid, float* ms, int32_t* best_grid, int32_t* best_first);

/* Measurement mode: time every launch of the pursuit update kernel with HIP events on the
 * handle's stream, and the persistent pursuit (when the window takes it) with its workgroups'
 * GPU-clock stamps; lss_last_update_stats gives the last window's summed update time and the
 * number of update launches (max_passes per window).  Results do not change. */
lss_status lss_profile_update(void* handle, int32_t enable);
lss_status lss_last_update_stats(void* handle, float* total_ms, int32_t* launches);
/* Per launch of the last window's update kernel (measurement mode): event_ms[i] between HIP events
 * recorded around it (includes the dispatch latency, ~9 us here, so short launches read long),
 * gpu_ms[i] from the workgroups' own GPU-clock stamps (last end - first start:
 * the execution time a kernel trace reports), spikes[i] the spikes it applied (0 for launches after
 * the pursuit converged), for i < min(n, *launches).  lss_profile_update must be enabled after the
 * update grid is set. */
lss_status lss_last_update_times(void* handle, float* event_ms, float* gpu_ms, int32_t* spikes, int32_t n,
                                 int32_t* launches);
/* Measurement mode, persistent pursuit (lss_set_persistent): for the last window, its GPU-clock
 * span (last workgroup end - first workgroup start), the spikes it applied, the passes it ran and
 * the workgroups it launched; all 0 when the window took the per-pass kernels only.  Its passes are
 * then absent from lss_last_update_times (their spikes read 0 there). */
lss_status lss_last_persistent_stats(void* handle, float* gpu_ms, int32_t* spikes, int32_t* passes, int32_t* grid);

/* ---- per-batch drift correction ----------------------------------------------------------------
 * Replaces the handle's whitening operator with (M @ Wrot), M a device [C][C] FP64 matrix (row
 * stride C, from lss_drift_matrix), the product taken in FP64 on the handle's stream and stored as
 * FP32; NULL restores Wrot alone.  Kilosort applies the batch's drift this way
 * (kilosort/io.py:781-787: M = get_drift_matrix(ops, dshift[ibatch]); X = (M @ Wrot) @ X); LSS
 * loads one fixed M (drift_matrix.npy, src/SpikeSorter/OnlineSpikesV2.cu:575-582).  Synchronous
 * (the operator is in place when it returns); in streaming mode call it between lss_stream_step
 * calls. */
lss_status lss_set_drift(void* handle, const double* d_m);

/* ---- training-side artifacts in FP64 -------------------------------------------------------------
 * Builders of the Kilosort4 artifacts the sorter loads, on the GPU in FP64.  Synchronous: each
 * call runs on hip_stream (NULL: the null stream) and waits for its work.  Scratch comes from a
 * per-device arena that grows to the largest request and is kept until lss_train_release (calls on
 * one device take turns on it).  d_* arguments are device memory of the current device; xc, yc,
 * yblk, dshift are host arrays. */
lss_status lss_train_release(void);

/* Covariance of one filtered batch, accumulated: cc[i][j] += sum_t x[i][t] x[j][t] / n over the
 * n columns of the [C][ldx] rows x (the loop body of get_whitening_matrix,
 * kilosort/preprocessing.py:96-113, which then divides by the number of batches).  Products and
 * sums in FP64, summation order fixed; cc comes out exactly symmetric. */
lss_status lss_covariance_add(const float* d_x, int64_t ldx, int32_t C, int64_t n, double* d_cc, int64_t ldcc,
                              void* hip_stream);
/* ZCA whitening W = E diag(1/sqrt(d + eps)) E^T of an n x n covariance, n <= 96
 * (whitening_from_covariance, kilosort/preprocessing.py:8-17; eps = 1e-6 there).  The
 * eigendecomposition is a parallel-order Jacobi in FP64; d enters as |d|, the singular values the
 * reference's SVD gives. */
lss_status lss_whitening_from_covariance(const double* d_cc, int64_t ldcc, int32_t n, double eps, double* d_w,
                                         int64_t ldw, void* hip_stream);
/* Local whitening (whitening_local, kilosort/preprocessing.py:19-35): for channel j the
 * min(nrange, C) <= 96 channels nearest to it (squared distance, ties by channel index), the ZCA
 * of cc over them, and row j of W = that ZCA's first row on those channels, zero elsewhere. */
lss_status lss_whitening_local(const double* d_cc, int64_t ldcc, int32_t C, const double* xc, const double* yc,
                               int32_t nrange, double eps, double* d_w, int64_t ldw, void* hip_stream);
/* iKxx = inv(Kxx + lam I), Kxx[i][j] = exp(-|p_i - p_j|^2 / (2 sig^2)) over the sites
 * p = (xc, yc) (kilosort/datashift.py:176-184, 215-219; lam = 0.01 there), by Newton-Schulz
 * iteration X <- X + X (I - A X) from X = I / ||A||_inf on the FP64 GEMM, stopped when the
 * residual ||I - A X||_F stops shrinking; iterations (may be NULL) gets the count. */
lss_status lss_drift_inverse(const double* xc, const double* yc, int32_t C, double sig_interp, double lam,
                             double* d_ikxx, int64_t ldk, int32_t* iterations, void* hip_stream);
/* Drift interpolation matrix M = Kyx @ iKxx (get_drift_matrix, kilosort/preprocessing.py:44-69):
 * the sites shifted down by the per-channel shift, which is dshift[0] for nblocks == 1 and
 * otherwise the linear interpolation (extrapolated at both ends) of dshift[b] over yblk[b] at yc. */
lss_status lss_drift_matrix(const double* xc, const double* yc, int32_t C, const double* yblk, const double* dshift,
                            int32_t nblocks, double sig_interp, const double* d_ikxx, int64_t ldk, double* d_m,
                            int64_t ldm, void* hip_stream);
/* Isolated single-channel peaks and their snippets (extract_snippets,
 * kilosort/spikedetect.py:32-47) of a whitened time-major window x [W][ldx]: a peak is a sample
 * whose |x| is the maximum over channels c +- loc_range[0] and samples t +- loc_range[1] and
 * exceeds th_single_ch; it is isolated when it is the only peak over c +- long_range[0],
 * t +- long_range[1]; samples t < nt and t >= W - nt are excluded.  Snippet i (of peak (c_i, t_i),
 * in channel-then-sample order) is clips[i][m] = x[t_i - twav_min + m][c_i], m < nt; the first
 * max_clips are written (chan, time may be NULL) and n_found gets the total.  loc_range <= 8. */
lss_status lss_extract_snippets(const float* d_x_wc, int64_t ldx, int32_t C, int64_t W, int32_t nt, int32_t twav_min,
                                float th_single_ch, const int32_t* loc_range, const int32_t* long_range,
                                float* d_clips, int64_t max_clips, int32_t* d_chan, int32_t* d_time, int64_t* n_found,
                                void* hip_stream);
/* wPCA [n_pcs][nt] (FP64) and singular values [n_pcs]: the leading right singular vectors of the
 * row-normalised snippets [n][nt] (TruncatedSVD(n_pcs).fit(clips / ||clips||) of
 * extract_wPCA_wTEMP, kilosort/spikedetect.py:50-72), each signed so that its largest-magnitude
 * entry is positive (the sign rule TruncatedSVD applies); nt <= 96. */
lss_status lss_wpca(const float* d_clips, int64_t n, int32_t nt, int32_t n_pcs, double* d_wpca, double* d_sv,
                    void* hip_stream);

/* Allocation-free lss_extract_snippets: same arguments, results and errors, for callers that
 * cannot allocate device memory (tight acquisition or serving loops, harnesses that forbid
 * hipMalloc).  It has no counterpart in the reference.  Its scratch (two peak bitmaps of
 * 8 * C * ceil(W / 64) bytes each and the scan's block sums) lives in a fixed device array of
 * LSS_NOALLOC_SCRATCH_BYTES on each GPU, held from the library's load until the process ends; no
 * call allocates or frees device memory.  A window whose scratch does not fit (C * ceil(W / 64)
 * above about 520000, e.g. 384 channels by 88000 samples) fails with LSS_ERR_ARG before any GPU
 * work.  Calls on one device share its array: they take turns on a lock held until their kernels
 * are done, and every word a call uses is written before it is read, so the array's leftover
 * contents never matter.  Synchronous, like lss_extract_snippets. */
#define LSS_NOALLOC_SCRATCH_BYTES (8u << 20)
lss_status lss_extract_snippets_noalloc(const float* d_x_wc, int64_t ldx, int32_t C, int64_t W, int32_t nt,
                                        int32_t twav_min, float th_single_ch, const int32_t* loc_range,
                                        const int32_t* long_range, float* d_clips, int64_t max_clips, int32_t* d_chan,
                                        int32_t* d_time, int64_t* n_found, void* hip_stream);

#ifdef __cplusplus
}
#endif
#endif
