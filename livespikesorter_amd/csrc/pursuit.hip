// Greedy matching pursuit over the score matrix B [W][ldT] (time-major: the T template scores of a
// sample are contiguous), Kilosort4 semantics (kilosort/template_matching.py:157-198):
//
//   Cf = relu(B)^2 / nm, zero in the first and last nt columns             (:159, :165-166)
//   Cfmax[w], imax[w] = max / first argmax over templates                  (:167)
//   peaks: maxpool_{2nt+1}(Cfmax)[w] > Th^2 and |pool - Cfmax[w]| < 1e-9   (:168-174)
//   amp = B[imax, w] / nm[imax]                                            (:188)
//   B[:, w-nt..w+nt] -= amp ctc[:, t, :];  Xres[:, w-nt/2..w+nt/2] -= amp wave_t
//   in two groups (even, then odd spike index of the pass)                 (:195-198)
//
// Scores are reduced with one wavefront per sample column (wavefront-shuffle argmax) in the first
// pass; afterwards every updated score column is rescored in place, so Cfmax/imax stay current
// without rescanning B.  Peak flags are computed over the whole chip and compacted in ascending
// column order by one workgroup.  The score update is a gather over (spike, column chunk) work
// items that streams ctc rows; the residual update is deferred to one kernel after the last pass
// (see the sections below for how both reproduce the reference's update order exactly).
//
// Spike counts live on the device: the pass loop is enqueued without host round trips; a pass
// after convergence finds no peaks and its update exits at once.
//
// FP contraction is off in this file: products and differences are rounded separately, as the
// reference's separate tensor multiply and in-place subtraction are.
#pragma clang fp contract(off)
#include "lss_kernels.h"

namespace lss {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned uint32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void argmax_merge(float& v, int& i, float v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

__device__ __forceinline__ void wave_argmax(float& v, int& i) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float v2 = __shfl_xor(v, off, 64);
        int i2 = __shfl_xor(i, off, 64);
        argmax_merge(v, i, v2, i2);
    }
}

// The same with the winner's raw score b carried along (the persistent pursuit publishes each
// rescored column's amplitude b / nm[i] beside its Cfmax/imax).
__device__ __forceinline__ void argmax_merge_b(float& v, int& i, float& b, float& m, float v2, int i2, float b2,
                                               float m2) {
    if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; b = b2; m = m2; }
}

// b = the winner's raw score B[w][i], m = its nm[i]
__device__ __forceinline__ void wave_argmax_b(float& v, int& i, float& b, float& m) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float v2 = __shfl_xor(v, off, 64);
        int i2 = __shfl_xor(i, off, 64);
        float b2 = __shfl_xor(b, off, 64);
        float m2 = __shfl_xor(m, off, 64);
        argmax_merge_b(v, i, b, m, v2, i2, b2, m2);
    }
}

// Wavefront maximum (uniform result): DPP steps within rows of 16 lanes, then the row broadcasts.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, v), __builtin_bit_cast(int, v),
                                                                 CTRL, ROWS, 0xF, false));
}
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ float wave_max(float v) {
    v = fmaxf(v, dpp_f<0xB1>(v));         // quad_perm [1, 0, 3, 2]
    v = fmaxf(v, dpp_f<0x4E>(v));         // quad_perm [2, 3, 0, 1]
    v = fmaxf(v, dpp_f<0x141>(v));        // row_half_mirror
    v = fmaxf(v, dpp_f<0x140>(v));        // row_mirror
    v = fmaxf(v, dpp_f<0x142, 0xA>(v));   // row_bcast:15 into rows 1 and 3
    v = fmaxf(v, dpp_f<0x143, 0xC>(v));   // row_bcast:31 into rows 2 and 3
    return readlane_f(v, 63);
}

// Wavefront inclusive prefix sum: shifts by 1, 2, 4, 8 within rows of 16 lanes (zero from outside
// the row), then the last lane of each row added into the rows after it.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ int dpp_i0(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false); }
__device__ __forceinline__ int wave_incl_scan(int v) {
    v += dpp_i0<0x111>(v);         // row_shr:1
    v += dpp_i0<0x112>(v);         // row_shr:2
    v += dpp_i0<0x114>(v);         // row_shr:4
    v += dpp_i0<0x118>(v);         // row_shr:8
    v += dpp_i0<0x142, 0xA>(v);    // row_bcast:15 into rows 1 and 3
    v += dpp_i0<0x143, 0xC>(v);    // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ float score_of(float b, float nmv) {
    float r = fmaxf(b, 0.f);
    return (r * r) / nmv;
}

// Column rescoring, exact: Cfmax = max_t q_t with q_t = fl(fl(r_t r_t) / nm_t), r_t = max(B[w][t], 0)
// (the reference's relu(B)**2 / nm), and its first argmax, over t < T; x[i][e] holds the column's
// B[w][4 lane + 256 i + e].  The division is only evaluated where it can decide the result: with
// s_t = fl(fl(r_t r_t) rn_t), rn_t = fl(1 / nm_t), |s_t / q_t - 1| < 2^-22 in the normal range, so
// every t with q_t = max q has s_t >= M (1 - 2^-20), M = max s.  Those candidates (nearly always a
// single one) get the exact quotient and the first-maximum merge; M outside [1e-30, 1e30] makes
// every template a candidate.  When every lane holds at most one candidate, each takes one
// quotient, its element selected without branches.  M = 0 with every fl(r_t r_t) = 0 (no positive
// score) gives q_t = 0 for all t: the first maximum is t = 0.  With DEV, b and m receive the
// winner's B[w][t] and nm_t.
constexpr float kCandidate = 0.99999905f;   // 1 - 2^-20
template <int NV, bool DEV>
__device__ __forceinline__ void rescore(const floatx4* x, const floatx4* nmv, const floatx4* rnv, int T, float& best,
                                        int& bi, float& bb, float& bm) {
    const int lane = threadIdx.x & 63;
    float sa[NV][4];
    float ml = -1.f, pl = 0.f;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float r = fmaxf(x[i][e], 0.f);
            const float p = r * r;
            const float sv = p * rnv[i][e];
            const bool in = 4 * lane + 256 * i + e < T;
            sa[i][e] = in ? sv : -1.f;
            ml = fmaxf(ml, sa[i][e]);
            pl = fmaxf(pl, in ? p : 0.f);
        }
    const float M = wave_max(ml);
    best = -1.f;
    bi = 0x7fffffff;
    bb = 0.f;
    bm = 1.f;
    if (M == 0.f && wave_max(pl) == 0.f) {
        // no positive score: every q_t = 0
        best = 0.f;
        bi = 0;
        if constexpr (DEV) {
            bb = readlane_f(x[0][0], 0);
            bm = readlane_f(nmv[0][0], 0);
        }
        return;
    }
    const float thr = (M >= 1e-30f && M <= 1e30f) ? M * kCandidate : -0.5f;
    uint32_t cm = 0;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) cm |= (sa[i][e] >= thr ? 1u : 0u) << (4 * i + e);
    if (__ballot(__popc(cm) > 1) == 0) {
        // at most one candidate per lane
        if (cm) {
            const int k = __ffs(cm) - 1;
            float xv = 0.f, nv = 1.f;
#pragma unroll
            for (int i = 0; i < NV; ++i)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (k == 4 * i + e) {
                        xv = x[i][e];
                        nv = nmv[i][e];
                    }
            best = score_of(xv, nv);
            bi = 4 * lane + 256 * (k >> 2) + (k & 3);
            if constexpr (DEV) {
                bb = xv;
                bm = nv;
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (sa[i][e] >= thr) {
                    const float q = score_of(x[i][e], nmv[i][e]);
                    if (q > best) {   // ascending t within the lane: the first maximum stays
                        best = q;
                        bi = 4 * lane + 256 * i + e;
                        if constexpr (DEV) {
                            bb = x[i][e];
                            bm = nmv[i][e];
                        }
                    }
                }
    }
    const unsigned long long cand = __ballot(bi != 0x7fffffff);
    if (__popcll(cand) == 1) {
        const int l = __ffsll((long long)cand) - 1;
        best = readlane_f(best, l);
        bi = __builtin_amdgcn_readlane(bi, l);
        if constexpr (DEV) {
            bb = readlane_f(bb, l);
            bm = readlane_f(bm, l);
        }
    } else if (cand != 0) {
        if constexpr (DEV) wave_argmax_b(best, bi, bb, bm);
        else wave_argmax(best, bi);
    }
}

// Wavefront maxima of NC values at once (uniform results), their DPP steps interleaved.  The values
// are non-negative or -1, whose bit patterns order as signed integers the way the floats do.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ int dpp_i(int v) { return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xF, false); }
template <int NC>
__device__ __forceinline__ void wave_max_n(float* v) {
    int x[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = __builtin_bit_cast(int, v[c]);
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = max(x[c], dpp_i<0xB1>(x[c]));         // quad_perm [1, 0, 3, 2]
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = max(x[c], dpp_i<0x4E>(x[c]));         // quad_perm [2, 3, 0, 1]
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = max(x[c], dpp_i<0x141>(x[c]));        // row_half_mirror
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = max(x[c], dpp_i<0x140>(x[c]));        // row_mirror
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = max(x[c], dpp_i<0x142, 0xA>(x[c]));   // row_bcast:15 into rows 1, 3
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = max(x[c], dpp_i<0x143, 0xC>(x[c]));   // row_bcast:31 into rows 2, 3
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(x[c], 63));
}

// rescore<NV, true> of NC columns x[c] (c < nc valid) with their steps interleaved: the same
// results column by column.
template <int NV, int NC>
__device__ __forceinline__ void rescore_n(const floatx4 (*x)[NV], const floatx4* nmv, const floatx4* rnv, int T, int nc,
                                          float* best, int* bi, float* bb, float* bm) {
    const int lane = threadIdx.x & 63;
    float sa[NC][NV][4], ml[NC], pl[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        ml[c] = -1.f;
        pl[c] = 0.f;
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float r = fmaxf(x[c][i][e], 0.f);
                const float p = r * r;
                const float sv = p * rnv[i][e];
                const bool in = 4 * lane + 256 * i + e < T;
                sa[c][i][e] = in ? sv : -1.f;
                ml[c] = fmaxf(ml[c], sa[c][i][e]);
                pl[c] = fmaxf(pl[c], in ? p : 0.f);
            }
    }
    wave_max_n<NC>(ml);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        best[c] = -1.f;
        bi[c] = 0x7fffffff;
        bb[c] = 0.f;
        bm[c] = 1.f;
        if (c >= nc) continue;
        const float M = ml[c];
        if (M == 0.f && wave_max(pl[c]) == 0.f) {
            // no positive score: every q_t = 0
            best[c] = 0.f;
            bi[c] = 0;
            bb[c] = readlane_f(x[c][0][0], 0);
            bm[c] = readlane_f(nmv[0][0], 0);
            continue;
        }
        const float thr = (M >= 1e-30f && M <= 1e30f) ? M * kCandidate : -0.5f;
        uint32_t cm = 0;
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) cm |= (sa[c][i][e] >= thr ? 1u : 0u) << (4 * i + e);
        float b = -1.f, xb = 0.f, nb = 1.f;
        int k_ = 0x7fffffff;
        if (__ballot(__popc(cm) > 1) == 0) {
            if (cm) {
                const int k = __ffs(cm) - 1;
#pragma unroll
                for (int i = 0; i < NV; ++i)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (k == 4 * i + e) {
                            xb = x[c][i][e];
                            nb = nmv[i][e];
                        }
                b = score_of(xb, nb);
                k_ = 4 * lane + 256 * (k >> 2) + (k & 3);
            }
        } else {
#pragma unroll
            for (int i = 0; i < NV; ++i)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (sa[c][i][e] >= thr) {
                        const float q = score_of(x[c][i][e], nmv[i][e]);
                        if (q > b) {   // ascending t within the lane: the first maximum stays
                            b = q;
                            k_ = 4 * lane + 256 * i + e;
                            xb = x[c][i][e];
                            nb = nmv[i][e];
                        }
                    }
        }
        const unsigned long long cand = __ballot(k_ != 0x7fffffff);
        if (__popcll(cand) == 1) {
            const int l = __ffsll((long long)cand) - 1;
            b = readlane_f(b, l);
            k_ = __builtin_amdgcn_readlane(k_, l);
            xb = readlane_f(xb, l);
            nb = readlane_f(nb, l);
        } else if (cand != 0) {
            wave_argmax_b(b, k_, xb, nb);
        }
        best[c] = b;
        bi[c] = k_;
        bb[c] = xb;
        bm[c] = nb;
    }
}

// Device-scope (all XCDs) relaxed loads and stores: the only way workgroups of one launch of the
// persistent pursuit exchange data, so that its grid barrier needs no cache write-back/invalidate.
template <typename V>
__device__ __forceinline__ V ld_dev(const V* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename V>
__device__ __forceinline__ void st_dev(V* p, V v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// 16-byte loads of values other workgroups of the same launch stored with st_dev: buffer loads with
// the device-scope cache policy (the same one as the device-scope loads above), through a resource
// over `bytes` bytes at p (the hardware returns 0 past the end).
constexpr int kCoherent = 16;                   // buffer cache policy: device scope
constexpr int kRawBufferWord3 = 0x00020000;     // raw (untyped) buffer
__device__ __forceinline__ uint32x4 ld_dev16(const void* p, int bytes, int off) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, bytes, kRawBufferWord3);
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kCoherent);
}

// Cfmax/imax of column w: plain stores between kernels; inside the persistent pursuit (DEV)
// device-scope stores of Cfmax, imax and the amplitude b / m the column would have as a peak
// (b = B[w][i], m = nm[i]: the expression the per-pass compaction evaluates)
template <bool DEV>
__device__ __forceinline__ void store_score(const PursuitArgs& a, int64_t w, float v, int i, float b = 0.f,
                                            float m = 1.f) {
    if constexpr (DEV) {
        st_dev(a.cfmax + w, v);
        st_dev(a.imax + w, i);
        st_dev(a.amprow + w, b / m);
    } else {
        a.cfmax[w] = v;
        a.imax[w] = i;
    }
}

// ------------------------------------------------------------------------------------------------
// full scan (first pass): one wavefront per column, float4 loads (ldT % 4 == 0)
// ------------------------------------------------------------------------------------------------
template <int NV>   // float4 vectors per lane: NV*256 >= T
__device__ __forceinline__ void score_column(const PursuitArgs& a, int64_t w, int lane, const floatx4* nmv,
                                             const floatx4* rnv) {
    float best = 0.f, bb = 0.f, bm = 1.f;
    int bi = 0;
    if (w >= a.nt && w < a.W - a.nt) {
        floatx4 b[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int t = 4 * lane + 256 * i;
            b[i] = (t < a.T) ? *(const floatx4*)(a.B + w * a.ldT + t) : floatx4{0.f, 0.f, 0.f, 0.f};
        }
        rescore<NV, false>(b, nmv, rnv, a.T, best, bi, bb, bm);
    }
    if (lane == 0) store_score<false>(a, w, best, bi, bb, bm);
}

// template norms nm and their reciprocals rn in the lanes' float4 layout (1 past T)
template <int NV>
__device__ __forceinline__ void load_norms(const PursuitArgs& a, floatx4* nmv, floatx4* rnv) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int t = 4 * lane + 256 * i;
        nmv[i] = (t < a.T) ? *(const floatx4*)(a.nm + t) : floatx4{1.f, 1.f, 1.f, 1.f};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (t + e >= a.T) nmv[i][e] = 1.f;
            rnv[i][e] = 1.f / nmv[i][e];
        }
    }
}

// one wavefront per column, grid-stride over the columns
template <int NV>
__global__ __launch_bounds__(256) void score_all_kernel(PursuitArgs a) {
    floatx4 nmv[NV], rnv[NV];
    load_norms<NV>(a, nmv, rnv);
    for (int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < a.W; w += (int64_t)gridDim.x * 4)
        score_column<NV>(a, w, threadIdx.x & 63, nmv, rnv);
}

// ------------------------------------------------------------------------------------------------
// peaks.  A column is a peak iff pool > Th^2 and |pool - cf| < 1e-9 (pool = window max).  If any
// window value v has fl(v - cf) >= 1e-9 then fl(pool - cf) >= fl(v - cf) (rounding is monotone),
// so the column is not a peak: the scan stops at the first such v, nearest neighbours first.  A
// column with cf <= Th^2 - 2e-9 cannot reach pool > Th^2 with pool - cf < 1e-9 and is skipped.
// Columns that survive get the literal test on the full window maximum.
//
// peaks_kernel: flags over the whole chip, compaction by the last workgroup (below).
// ------------------------------------------------------------------------------------------------
constexpr int kFlagCols = 1024;
constexpr int kPeakHalo = 128;                       // >= nt

// p points at the column's Cfmax in a buffer holding its +-nt neighbours (-inf outside the window).
// The neighbours are read 8 distances at a time (all 16 loads in flight together); distances past
// nt read the distance-nt values again, which changes neither the test nor the maximum.
constexpr int kPeakStep = 8;
__device__ __forceinline__ bool peak_test(const float* p, float th2, int nt) {
    const float cf = p[0];
    if (!((double)cf > (double)th2 - 2e-9) || p[-1] - cf >= 1e-9f || p[1] - cf >= 1e-9f) return false;
    float cm = fmaxf(cf, fmaxf(p[-1], p[1]));
    for (int d0 = 2; d0 <= nt; d0 += kPeakStep) {
        float vl[kPeakStep], vr[kPeakStep];
#pragma unroll
        for (int j = 0; j < kPeakStep; ++j) {
            const int d = min(d0 + j, nt);
            vl[j] = p[-d];
            vr[j] = p[d];
        }
        bool bad = false;
#pragma unroll
        for (int j = 0; j < kPeakStep; ++j) {
            bad |= (vl[j] - cf >= 1e-9f) | (vr[j] - cf >= 1e-9f);
            cm = fmaxf(cm, fmaxf(vl[j], vr[j]));
        }
        if (bad) return false;
    }
    return cm > th2 && fabsf(cm - cf) < 1e-9f;
}

// peaks_kernel: one thread per column over the whole chip; each wavefront publishes its 64 peak
// flags as two 32-bit words (device-scope stores), takes a ticket, and the last workgroup to finish
// compacts the flag words in ascending column order (one launch per pass, no waiting).
constexpr int kEmitPer = 64;                          // columns per thread per compaction tile
constexpr int kEmitTile = kFlagCols * kEmitPer;       // 65536

__device__ __forceinline__ void emit_pass(const PursuitArgs& a, const uint32_t* flagw) {
    __shared__ int wsum[kFlagCols / 64];
    __shared__ int s_base, s_run;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t nwords = (a.W + 31) / 32;
    // the first tile's flag words, loaded while thread 0 reads the counters
    const uint32_t lo0 = (2 * tid < nwords) ? ld_dev(flagw + 2 * tid) : 0u;
    const uint32_t hi0 = (2 * tid + 1 < nwords) ? ld_dev(flagw + 2 * tid + 1) : 0u;
    if (tid == 0) {
        // commit the previous pass (its update has completed: stream order)
        const int base = a.counters[0] + a.counters[1];
        a.counters[0] = base;
        a.counters[1] = 0;
        s_base = base;
        s_run = 0;
    }
    __syncthreads();
    for (int64_t c0 = 0; c0 < a.W; c0 += kEmitTile) {
        const int64_t wd = (c0 >> 5) + 2 * tid;
        const uint32_t lo = (c0 == 0) ? lo0 : (wd < nwords) ? ld_dev(flagw + wd) : 0u;
        const uint32_t hi = (c0 == 0) ? hi0 : (wd + 1 < nwords) ? ld_dev(flagw + wd + 1) : 0u;
        uint64_t bits = ((uint64_t)hi << 32) | lo;
        const int cnt = __popcll(bits);
        int inc = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int vv = __shfl_up(inc, off, 64);
            if (lane >= off) inc += vv;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int woff = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < kFlagCols / 64; ++k) {
            const int vv = wsum[k];
            if (k < wave) woff += vv;
            tot += vv;
        }
        int pos = s_base + s_run + woff + inc - cnt;
        while (bits) {
            const int k = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            const int64_t w = c0 + (int64_t)tid * kEmitPer + k;
            if (pos < a.cap && w < a.W) {
                const int t = a.imax[w];
                a.sp_col[pos] = (int)w;
                a.sp_tmp[pos] = t;
                a.sp_amp[pos] = a.B[w * a.ldT + t] / a.nm[t];
            }
            ++pos;
        }
        __syncthreads();
        if (tid == 0) s_run += tot;
        __syncthreads();
    }
    if (tid == 0) {
        const int n = s_run;
        a.counters[1] = n;
        if (s_base + n > a.cap) a.counters[2] = 1;
        if (n == 0) {
            a.counters[3] = 1;
        } else {
            const int p = a.counters[4];
            a.pass_off[p] = s_base;
            a.pass_off[p + 1] = min(s_base + n, a.cap);
            a.counters[4] = p + 1;
        }
    }
}

__global__ __launch_bounds__(kFlagCols) void peaks_kernel(PursuitArgs a) {
    __shared__ float lds[kFlagCols + 2 * kPeakHalo];
    __shared__ int s_last;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t* flagw = (uint32_t*)a.flags;
    const int64_t c0 = (int64_t)blockIdx.x * kFlagCols;
    constexpr int kRows = (kFlagCols + 2 * kPeakHalo + kFlagCols - 1) / kFlagCols;
    // the row's loads go out together with the converged-flag load
    float v[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        const int64_t w = c0 - kPeakHalo + tid + r * kFlagCols;
        v[r] = (w >= 0 && w < a.W) ? a.cfmax[w] : -INFINITY;
    }
    if (a.counters[3]) return;   // converged earlier: nothing can change (no workgroup has ticketed)
#pragma unroll
    for (int r = 0; r < kRows; ++r)
        if (tid + r * kFlagCols < kFlagCols + 2 * kPeakHalo) lds[tid + r * kFlagCols] = v[r];
    __syncthreads();
    const int64_t w = c0 + tid;
    const bool pk = w < a.W && peak_test(lds + kPeakHalo + tid, a.th2, a.nt);
    const uint64_t bal = __ballot(pk);
    if (lane == 0) {
        st_dev(flagw + (c0 >> 5) + 2 * wave, (uint32_t)bal);
        st_dev(flagw + (c0 >> 5) + 2 * wave + 1, (uint32_t)(bal >> 32));
    }
    // ticket after this workgroup's words are stored; the last workgroup sees every word
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) s_last = atomicAdd(a.counters + 8, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (tid == 0) a.counters[8] = 0;   // ticket of the next pass
    emit_pass(a, flagw);
}

// ------------------------------------------------------------------------------------------------
// update: work items (spike s, chunk of 4 * CPW of its window's score columns, CPW columns per
// wavefront: 2 (fewer items, for passes with many spikes) or 1 (half the work per lane, for
// passes with few spikes); every column gets the same arithmetic either way)
// ------------------------------------------------------------------------------------------------
constexpr int kUpdateGrid = 8192;   // workgroups per update launch (grid-stride over work items)

// A pass's spike list (ascending columns), stored by the peak kernel.
struct ListPlain {
    const int* col;
    const int* tmp;
    const float* amp;
    __device__ int c(int i) const { return col[i]; }
    __device__ int t(int i) const { return tmp[i]; }
    __device__ float a(int i) const { return amp[i]; }
};

// Spikes of the pass covering column w within +-reach, starting at spike s (whose window covers
// w): the last even-indexed (e) and odd-indexed (o) ones.  Spikes are in ascending column order,
// so the covering ones are s, s+1, ...; c1, c2 are col[s+1], col[s+2] (INT_MAX past the end).
__device__ __forceinline__ void covering(const ListPlain& l, int s, int n, int c1, int c2, int w, int reach,
                                         int& e, int& o) {
    int last = s;
    if (c1 - reach <= w) {
        last = s + 1;
        if (c2 - reach <= w) {
            last = s + 2;
            while (last + 1 < n && l.c(last + 1) - reach <= w) ++last;   // tied peaks only
        }
    }
    if (last == s) {
        e = (s & 1) ? -1 : s;
        o = (s & 1) ? s : -1;
    } else {
        // the last two covering indices have different parities
        e = (last & 1) ? last - 1 : last;
        o = (last & 1) ? last : last - 1;
    }
}

// Columns wa and wb of this wavefront (va / vb: updated at all; sa / sb: the first spike of the
// pass whose window covers the column, with ca1, ca2 / cb1, cb2 the columns of the next two
// spikes): B -= amp ctc for the last even and last odd covering spike, then the column rescored.
template <int NV>   // float4 vectors per lane over templates: NV*256 >= T
__device__ __forceinline__ void update_two(const PursuitArgs& a, const ListPlain& l, int n, int wa, int sa, int ca1,
                                           int ca2, bool va, int wb, int sb, int cb1, int cb2, bool vb,
                                           const floatx4* nmv, const floatx4* rnv) {
    const int lane = threadIdx.x & 63;
    const int nt = a.nt, L = 2 * nt + 1;
    // two columns per wavefront, loads of both issued before any use
    int ea = -1, oa = -1, eb = -1, ob = -1;
    if (va) covering(l, sa, n, ca1, ca2, wa, nt, ea, oa);
    if (vb) covering(l, sb, n, cb1, cb2, wb, nt, eb, ob);
    floatx4 ba[NV], bb[NV], cea[NV], coa[NV], ceb[NV], cob[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) ba[i] = bb[i] = floatx4{0.f, 0.f, 0.f, 0.f};
    const float* qea = (ea >= 0) ? a.ctc + ((int64_t)l.t(ea) * L + (wa - l.c(ea) + nt)) * a.ldT : nullptr;
    const float* qoa = (oa >= 0) ? a.ctc + ((int64_t)l.t(oa) * L + (wa - l.c(oa) + nt)) * a.ldT : nullptr;
    const float* qeb = (eb >= 0) ? a.ctc + ((int64_t)l.t(eb) * L + (wb - l.c(eb) + nt)) * a.ldT : nullptr;
    const float* qob = (ob >= 0) ? a.ctc + ((int64_t)l.t(ob) * L + (wb - l.c(ob) + nt)) * a.ldT : nullptr;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int t = 4 * lane + 256 * i;
        if (t < a.T) {
            if (va) {
                ba[i] = *(const floatx4*)(a.B + (int64_t)wa * a.ldT + t);
                if (qea) cea[i] = *(const floatx4*)(qea + t);
                if (qoa) coa[i] = *(const floatx4*)(qoa + t);
            }
            if (vb) {
                bb[i] = *(const floatx4*)(a.B + (int64_t)wb * a.ldT + t);
                if (qeb) ceb[i] = *(const floatx4*)(qeb + t);
                if (qob) cob[i] = *(const floatx4*)(qob + t);
            }
        }
    }
    const float aea = (ea >= 0) ? l.a(ea) : 0.f, aoa = (oa >= 0) ? l.a(oa) : 0.f;
    const float aeb = (eb >= 0) ? l.a(eb) : 0.f, aob = (ob >= 0) ? l.a(ob) : 0.f;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int t = 4 * lane + 256 * i;
        if (va) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float v = ba[i][e];
                if (qea) v = v - aea * cea[i][e];
                if (qoa) v = v - aoa * coa[i][e];
                ba[i][e] = v;
            }
            if (t < a.T) *(floatx4*)(a.B + (int64_t)wa * a.ldT + t) = ba[i];
        }
        if (vb) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float v = bb[i][e];
                if (qeb) v = v - aeb * ceb[i][e];
                if (qob) v = v - aob * cob[i][e];
                bb[i][e] = v;
            }
            if (t < a.T) *(floatx4*)(a.B + (int64_t)wb * a.ldT + t) = bb[i];
        }
    }
    // rescored inside [nt, W - nt) only (the reference's Cf there; the edges never hold a peak)
    if (va && wa >= nt && wa < a.W - nt) {
        float best, b_, m_;
        int bi;
        rescore<NV, false>(ba, nmv, rnv, a.T, best, bi, b_, m_);
        if (lane == 0) store_score<false>(a, wa, best, bi, b_, m_);
    }
    if (vb && wb >= nt && wb < a.W - nt) {
        float best, b_, m_;
        int bi;
        rescore<NV, false>(bb, nmv, rnv, a.T, best, bi, b_, m_);
        if (lane == 0) store_score<false>(a, wb, best, bi, b_, m_);
    }
}

// Per-pass update: the chunk [c0, c0 + 4 CPW) of the columns [lo, w1] whose first covering spike is
// s (c1, c2 = the next two spikes' columns), CPW columns per wavefront.
template <int NV, int CPW>
__device__ __forceinline__ void update_score_columns(const PursuitArgs& a, const ListPlain& l, int s, int n, int c1, int c2,
                                                     int c0, int lo, int w1, const floatx4* nmv,
                                                     const floatx4* rnv) {
    const int wa = c0 + CPW * (int)(threadIdx.x >> 6), wb = wa + 1;
    update_two<NV>(a, l, n, wa, s, c1, c2, wa >= lo && wa <= w1, wb, s, c1, c2,
                              CPW == 2 && wb >= lo && wb <= w1, nmv, rnv);
}


// Work items q = q0, q0 + stride, ... of a pass's n spikes (col/tmp/amp, ascending columns; global
// or LDS memory).  A column is written only by the item of the first spike whose window covers it.
template <int NV, int CPW>
__device__ __forceinline__ void update_items(const PursuitArgs& a, const ListPlain& l, int n, int64_t q0,
                                             int64_t stride, const floatx4* nmv, const floatx4* rnv) {
    constexpr int kCols = 4 * CPW;
    const int nt = a.nt;
    const int nB = (2 * nt + 1 + kCols - 1) / kCols;
    for (int64_t q = q0; q < (int64_t)n * nB; q += stride) {
        const int s = (int)(q / nB), sub = (int)(q % nB);
        // the spike's column and its neighbours' (independent loads, issued together)
        const int ws = l.c(s);
        const int cp = (s > 0) ? l.c(s - 1) : INT_MIN / 2;
        const int c1 = (s + 1 < n) ? l.c(s + 1) : INT_MAX / 2;
        const int c2 = (s + 2 < n) ? l.c(s + 2) : INT_MAX / 2;
        int lo = ws - nt;
        if (s > 0) lo = max(lo, cp + nt + 1);
        lo = max(lo, 0);
        const int hi = (int)min((int64_t)ws + nt, a.W - 1);
        const int c0 = ws - nt + sub * kCols;
        const int w1 = min(c0 + kCols - 1, hi);
        if (max(c0, lo) <= w1) update_score_columns<NV, CPW>(a, l, s, n, c1, c2, c0, lo, w1, nmv, rnv);
    }
}

template <int NV, int CPW>
__global__ __launch_bounds__(256) void update_kernel(PursuitArgs a) {
    if (a.ts && threadIdx.x == 0) a.ts[2 * blockIdx.x] = __builtin_amdgcn_s_memrealtime();
    const int base = a.counters[0];
    int n = a.counters[1];
    if (base + n > a.cap) n = a.cap - base;
    if (n > 0) {
        floatx4 nmv[NV], rnv[NV];
        load_norms<NV>(a, nmv, rnv);
        update_items<NV, CPW>(a, ListPlain{a.sp_col + base, a.sp_tmp + base, a.sp_amp + base}, n, blockIdx.x,
                              gridDim.x, nmv, rnv);
    }
    if (a.ts) {
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's stores issued and done
        __syncthreads();
        if (threadIdx.x == 0) a.ts[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
    }
}

// ------------------------------------------------------------------------------------------------
// persistent pursuit for short windows (W <= kPersistMaxW, the streaming batches): the first
// scoring and every pass in one launch of G co-resident 512-thread workgroups, ONE grid barrier
// per pass, each workgroup's score rows resident in its LDS.
//
// Workgroup g owns the columns [a0, a1), a0 = gW/G (a fixed partition).  Their B rows are
// contiguous in B: they are copied into LDS at the start, updated there in every pass and copied
// back at the end, so the pass loop reads no score row from HBM; the workgroup alone scores them.
// What workgroups exchange goes through device-scope loads and stores only (no cache write-back or
// invalidate at the barrier): per column its Cfmax, imax and the amplitude B[w][imax]/nm[imax] it
// would have as a peak, in one of two rows chosen by pass parity.
//
// Pass p: every workgroup loads the Cfmax row p&1 of the whole window into LDS (with -inf halos),
// and imax/amplitude of its local range [a0 - nt, a1 + nt); it tests every column (peak_bits8) and
// counts the peaks before each 8-column block, so it knows the index of every peak in the pass's
// ascending list without exchanging it.  Its local list holds the peaks of the local range — the
// spikes whose window [col - nt, col + nt] can reach its columns — and their indices' parity: a
// column takes the last even-indexed, then the last odd-indexed covering spike
// (template_matching.py:195-198).  The owner of a peak's column stores it in the spike list.  Each
// workgroup then subtracts amp * ctc rows (streamed from HBM) from its covered columns' LDS rows,
// rescores them and publishes every own column into row (p+1)&1; the barrier ends the pass.  A
// workgroup writes row (p+1)&1 while others may still read row p&1, never the row they read, so the
// one barrier per pass orders every read of a pass before the writes that overwrite it.  Every
// workgroup sees the same row, so all take the same branches and leave the loop on the same pass.
//
// The barrier waits a bounded time: on a timeout (workgroups not co-resident) the kernel sets
// counters[7], every workgroup leaves, and the host reports the error; the launcher runs only grids
// the device holds at once with the LDS they need, so that this does not happen in normal operation.
// ------------------------------------------------------------------------------------------------
constexpr int kPT = 512;                                // threads per persistent workgroup
constexpr int kPW = kPT / 64;                           // its wavefronts
// LDS index of row entry i with one pad word per 32 (stride-8 reads conflict-free); pad_len(n) covers [0, n)
__device__ __forceinline__ int pad32(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int pad_len(int n) { return n + (n >> 5) + 1; }
constexpr int kBarrierSpin = 1 << 20;                   // ~1 s of polling
constexpr int kPersistLoads = (kPersistMaxW + 2 * kPeakHalo + kPT - 1) / kPT;   // row values per thread

// Grid barrier number `gen` (1, 2, ...) of this launch: each workgroup, once its memory operations
// have completed, stores base + gen into its flag bar[kBarStride g]; wavefront 0 then polls every
// workgroup's flag (device-scope loads, 64 per instruction) until all have reached base + gen.  No
// atomics, so arrivals do not queue on one address, and one flag per 64 bytes, so the pollers'
// loads spread over as many cache lines as there are workgroups.  base = the launch's epoch (counters[kEpoch], advanced by
// the counter reset before every pursuit) times 4096, so flags of earlier launches stay below it
// (differences taken modulo 2^32).  Data shared between workgroups is only accessed with
// device-scope operations, so no cache maintenance is needed.  Returns false after a timeout or
// when another workgroup has timed out.  g, G: this workgroup and the workgroups of its pursuit.
__device__ __forceinline__ bool grid_barrier(const PursuitArgs& a, unsigned base, int gen, int* s_ok, int g, int G) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const unsigned target = base + (unsigned)gen;
    unsigned* const bar = (unsigned*)a.bar;
    if (threadIdx.x == 0) st_dev(bar + kBarStride * g, target);
    if (threadIdx.x < 64) {
        int ok = 1;
        for (int it = 0;; ++it) {
            bool behind = false;
            for (int i = (int)threadIdx.x; i < G; i += 64)
                behind |= (int)(ld_dev(bar + kBarStride * i) - target) < 0;
            if (__ballot(behind) == 0) break;
            if (it >= kBarrierSpin || __ballot(ld_dev(a.counters + 7) != 0) != 0) {
                if (threadIdx.x == 0) atomicExch(a.counters + 7, 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (threadIdx.x == 0) *s_ok = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return *s_ok != 0;
}

// the row of pass parity `par`
__device__ __forceinline__ void use_row(PursuitArgs& b, const PursuitArgs& a, int par) {
    b.cfmax = a.prow_cf + (int64_t)par * kPersistMaxW;
    b.imax = a.prow_im + (int64_t)par * kPersistMaxW;
    b.amprow = a.prow_amp + (int64_t)par * kPersistMaxW;
}

// The persistent pursuit's peak test of one thread's 8 columns w = 8t + q (q < 8, w < W): the
// window maximum pool_q over [w - nt, w + nt] (-inf outside the row) and the reference's test
// pool > Th^2, |pool - cf| < 1e-9 (max-pool equality), as bits q of the result.  The row cf is
// in LDS at pad(H + column) (one pad word per 32: the stride-8 reads below are conflict-free) and
// bm[b + kBmHalo] holds the maximum of block b (columns 8b .. 8b + 7).  For nt >= 7 every window of
// the 8 columns contains C = [8t + 7 - nt, 8t + nt]: its maximum comes from the whole blocks inside
// it and the two partial runs at its ends; column q adds head[q .. 6] (head[k] = column
// 8t - nt + k) and tail[0 .. q - 1] (tail[k] = column 8t + nt + 1 + k).  Smaller nt scans directly.
// Every load is unconditional (so all are in flight together): indices past a run's end read the
// run's last entry again, and an empty partial run reads a column of the nearest whole block, both
// inside C, which changes no maximum.
constexpr int kBmHalo = 17;   // blocks of -inf kept on each side of the row (nt <= 127)
__device__ __forceinline__ uint32_t peak_bits8(const float* cf, const float* bm, int t, int W, int nt, float th2) {
    const int H = kPeakHalo, c0 = 8 * t;
    float pool[8];
    if (nt >= 7) {
        const int bl = -((nt - 7) >> 3), bh = ((nt + 1) >> 3) - 1;
        const int ph = (nt - 7) & 7, pt = (nt + 1) & 7;
        float cm = -INFINITY;
#pragma unroll
        for (int j = 0; j < 16; ++j) cm = fmaxf(cm, bm[min(t + bl + j, t + bh) + kBmHalo]);
        for (int b = t + bl + 16; b <= t + bh; ++b) cm = fmaxf(cm, bm[b + kBmHalo]);   // nt > 63
        const int s0 = c0 + 7 - nt, e1 = 8 * (t + bh + 1);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            cm = fmaxf(cm, cf[pad32(H + s0 + min(k, max(ph - 1, 0)))]);
            cm = fmaxf(cm, cf[pad32(H + e1 + min(k, pt - 1))]);
        }
        float hv[7], tv[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            hv[k] = cf[pad32(H + c0 - nt + k)];
            tv[k] = cf[pad32(H + c0 + nt + 1 + k)];
        }
        float suf = -INFINITY, pre = -INFINITY;
        float hs[8];
        hs[7] = -INFINITY;
#pragma unroll
        for (int k = 6; k >= 0; --k) {
            suf = fmaxf(suf, hv[k]);
            hs[k] = suf;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            pool[q] = fmaxf(cm, fmaxf(hs[q], pre));
            if (q < 7) pre = fmaxf(pre, tv[q]);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            float m = -INFINITY;
            for (int d = -nt; d <= nt; ++d) m = fmaxf(m, cf[pad32(H + c0 + q + d)]);
            pool[q] = m;
        }
    }
    uint32_t bits = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float c = cf[pad32(H + c0 + q)];
        if (c0 + q < W && pool[q] > th2 && fabsf(pool[q] - c) < 1e-9f) bits |= 1u << q;
    }
    return bits;
}

// LDS of the persistent pursuit (4-byte words) for W columns, row stride ldT and G workgroups: own
// B rows [ceil(W/G)][ldT] | template norms and their reciprocals [ldT] each | the padded Cfmax row
// with its halos and a -inf tail of 16 | block maxima with their halos | the row's imax and
// amplitudes [W] | the pass's list: column | template << 16, amplitude [W] | per own column the last
// even- and odd-indexed spike covering it | the covered own columns: own index, those two spikes
// [ceil(W/G)] each
__host__ __device__ constexpr int persist_cf_len(int W) { return pad_len(W + 2 * kPeakHalo + 16); }
__host__ __device__ constexpr int persist_bm_len(int W) { return (W + 7) / 8 + 2 * kBmHalo; }
struct PersistLayout {
    int own, nm, rn, cf, bm, rim, ram, lct, lamp, cove, covo, cvj, cvl, cvp, words;
    __host__ __device__ PersistLayout(int W, int ldT, int G) {
        own = (W + G - 1) / G;
        nm = own * ldT;
        rn = nm + ldT;
        cf = rn + ldT;
        bm = cf + persist_cf_len(W);
        rim = bm + persist_bm_len(W);
        ram = rim + W;
        lct = ram + W;
        lamp = lct + W;
        cove = lamp + W;
        covo = cove + own;
        cvj = covo + own;
        cvl = cvj + own;
        cvp = cvl + own;
        words = cvp + own;
    }
};

// floor(x / d) for 0 <= x < 2^24, 1 <= d, r = 1/d in FP32: the float quotient is off by at most one
__device__ __forceinline__ int floor_div(int x, int d, float r) {
    int q = (int)((float)x * r);
    if ((q + 1) * d <= x) ++q;
    if (q * d > x) --q;
    return q;
}

// Up to NC covered own columns of one wavefront (n2 <= NC of them): own index j, column w, and the
// last even- and last odd-indexed spikes of the pass whose window covers it (se / so, list index =
// index in the pass; -1 where none).  The LDS row gets B -= amp ctc for the even-indexed spike, then
// for the odd-indexed one (ctc rows loaded straight into registers, every load of the columns in
// flight at once), and is rescored into row nx inside [nt, W - nt) (columns nearer an edge are
// carried by the caller).
template <int NV, int NC>
__device__ __forceinline__ void persist_update_pair(const PursuitArgs& a, const PursuitArgs& nx, floatx4* Bs4, int ld4,
                                                    const int* lct, const float* lamp, int n2, const int* j,
                                                    const int* w, const int* se, const int* so, const floatx4* nm4,
                                                    const floatx4* rn4) {
    const int lane = threadIdx.x & 63;
    const int nt = a.nt, L = 2 * nt + 1;
    const floatx4 z = {0.f, 0.f, 0.f, 0.f};
    const float* r1[NC];
    const float* r2[NC];
    float a1[NC], a2[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int cc = c < n2 ? c : 0;
        const int e = se[cc], o = so[cc];
        r1[c] = e >= 0 ? a.ctc + ((int64_t)(lct[e] >> 16) * L + (w[cc] - (lct[e] & 0xFFFF) + nt)) * a.ldT : nullptr;
        r2[c] = o >= 0 ? a.ctc + ((int64_t)(lct[o] >> 16) * L + (w[cc] - (lct[o] & 0xFFFF) + nt)) * a.ldT : nullptr;
        a1[c] = e >= 0 ? lamp[e] : 0.f;
        a2[c] = o >= 0 ? lamp[o] : 0.f;
    }
    floatx4 x[NC][NV], c1[NC][NV], c2[NC][NV];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int t = 4 * lane + 256 * i;
            x[c][i] = c1[c][i] = c2[c][i] = z;
            if (c < n2 && t < a.T) {
                if (r1[c]) c1[c][i] = *(const floatx4*)(r1[c] + t);
                if (r2[c]) c2[c][i] = *(const floatx4*)(r2[c] + t);
                x[c][i] = Bs4[j[c] * ld4 + lane + 64 * i];
            }
        }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            floatx4 v = x[c][i];
            if (r1[c]) {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = v[e] - a1[c] * c1[c][i][e];
            }
            if (r2[c]) {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = v[e] - a2[c] * c2[c][i][e];
            }
            x[c][i] = v;
            if (c < n2 && 4 * lane + 256 * i < a.T) Bs4[j[c] * ld4 + lane + 64 * i] = v;
        }
    // rescored (every column; published inside [nt, W - nt))
    floatx4 nmv[NV], rnv[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const bool in = lane + 64 * i < ld4;
        nmv[i] = in ? nm4[lane + 64 * i] : floatx4{1.f, 1.f, 1.f, 1.f};
        rnv[i] = in ? rn4[lane + 64 * i] : floatx4{1.f, 1.f, 1.f, 1.f};
    }
    float best[NC], b_[NC], m_[NC];
    int bi[NC];
    rescore_n<NV, NC>(x, nmv, rnv, a.T, n2, best, bi, b_, m_);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c < n2 && w[c] >= nt && w[c] < a.W - nt) store_score<true>(nx, w[c], best[c], bi[c], b_[c], m_[c]);
}

// The persistent pursuit of one window by workgroup g of the G that run it (a launch of its own, or
// a share of a grouped launch).
template <int NV>
__device__ __forceinline__ void persistent_body(const PursuitArgs& a, int first_pass, int max_passes, int g, int G) {
    extern __shared__ floatx4 psm4[];
    float* const psm = (float*)psm4;
    __shared__ int wsum[kPW];
    __shared__ int s_ok;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = (int)a.W, H = kPeakHalo, nt = a.nt, L = 2 * nt + 1;
    const int N = W + 2 * H;
    const int nbk = (W + 7) / 8;   // blocks of 8 columns, one per thread
    const PersistLayout ly(W, (int)a.ldT, G);
    float* cf = psm + ly.cf;
    float* bm = psm + ly.bm;
    int* lct = (int*)(psm + ly.lct);
    float* lamp = psm + ly.lamp;
    int* rim = (int*)(psm + ly.rim);
    float* ram = psm + ly.ram;
    int* cove = (int*)(psm + ly.cove);
    int* covo = (int*)(psm + ly.covo);
    int* cvj = (int*)(psm + ly.cvj);
    int* cvl = (int*)(psm + ly.cvl);
    int* cvp = (int*)(psm + ly.cvp);
    const bool lead = g == 0 && tid == 0;
    if (a.ts && tid == 0) a.ts[2 * g] = __builtin_amdgcn_s_memrealtime();   // measurement mode
    // own columns g + G j (j < nown): strided, so that every workgroup holds a share of the columns
    // a pass updates wherever its spikes are; their B rows into LDS
    const int nown = g < W ? (W - 1 - g) / G + 1 : 0;
    const float rG = 1.f / (float)G;
    const int ld4 = (int)(a.ldT / 4);
    for (int j = wave; j < nown; j += kPW) {
        const floatx4* src = (const floatx4*)(a.B + (int64_t)(g + G * j) * a.ldT);
        for (int i = lane; i < ld4; i += 64) psm4[j * ld4 + i] = src[i];
    }
    // template norms and reciprocals in the lanes' float4 layout (1 past T)
    floatx4* const nm4 = psm4 + ly.nm / 4;
    floatx4* const rn4 = psm4 + ly.rn / 4;
    if (wave == 0) {
        floatx4 nmv[NV], rnv[NV];
        load_norms<NV>(a, nmv, rnv);
#pragma unroll
        for (int i = 0; i < NV; ++i)
            if (lane + 64 * i < ld4) {
                nm4[lane + 64 * i] = nmv[i];
                rn4[lane + 64 * i] = rnv[i];
            }
    }
    int par = first_pass & 1;
    PursuitArgs cur = a;
    use_row(cur, a, par);
    int base = 0, np = 0;
    bool done = false;
    // constant -inf parts: the row's halos and tail, the block halos
    for (int i = tid; i < H; i += kPT) cf[pad32(i)] = -INFINITY;
    for (int i = H + 8 * nbk + tid; i < N + 16; i += kPT) cf[pad32(i)] = -INFINITY;
    if (tid < kBmHalo) {
        bm[tid] = -INFINITY;
        bm[kBmHalo + nbk + tid] = -INFINITY;
    }
    __syncthreads();
    if (first_pass == 0) {
        // first pass: the own columns scored from LDS, one wavefront per column
        for (int j = wave; j < nown; j += kPW) {
            const int w = g + G * j;
            float best = 0.f, bb = 0.f, bmv = 1.f;
            int bi = 0;
            if (w >= nt && w < W - nt) {
                const floatx4* r = psm4 + j * ld4;
                floatx4 x[NV], nmv[NV], rnv[NV];
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    const bool in = lane + 64 * i < ld4;
                    x[i] = (4 * lane + 256 * i < a.T) ? r[lane + 64 * i] : floatx4{0.f, 0.f, 0.f, 0.f};
                    nmv[i] = in ? nm4[lane + 64 * i] : floatx4{1.f, 1.f, 1.f, 1.f};
                    rnv[i] = in ? rn4[lane + 64 * i] : floatx4{1.f, 1.f, 1.f, 1.f};
                }
                rescore<NV, true>(x, nmv, rnv, a.T, best, bi, bb, bmv);
            }
            if (lane == 0) store_score<true>(cur, w, best, bi, bb, bmv);
        }
    } else {
        // after first_pass passes of the per-pass kernels (their last pass committed as the peak
        // kernel commits it): own columns' Cfmax/imax as they left them, with their amplitudes
        base = a.counters[0] + a.counters[1];
        np = a.counters[4];
        done = a.counters[3] != 0;
        for (int j = tid; j < nown; j += kPT) {
            const int w = g + G * j, i = a.imax[w];
            store_score<true>(cur, w, a.cfmax[w], i, psm[(int64_t)j * a.ldT + i], a.nm[i]);
        }
    }
    const unsigned bbase = (unsigned)a.counters[kEpoch] * 4096u;
    int gen = 1;
    if (!grid_barrier(a, bbase, gen++, &s_ok, g, G)) return;
    // this thread's block: columns c0 .. c0 + 7, r0 = c0 mod G (the first is owned when r0 == g)
    const int c0 = 8 * tid, r0 = c0 % G;
    for (int pass = first_pass; pass < max_passes && !done; ++pass) {
        // the row: each thread its block's Cfmax (with the block maximum), imax and amplitudes into
        // LDS, every load in flight at once
        if (tid < nbk) {
            float cv[8], av[8];
            int iv[8];
            floatx4 fc[2], fa[2];
            uint32x4 ui[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                // whole blocks stay inside the row's kPersistMaxW entries (8 ceil(W/8) <= kPersistMaxW)
                fc[h] = __builtin_bit_cast(floatx4, ld_dev16(cur.cfmax, 4 * kPersistMaxW, 4 * (c0 + 4 * h)));
                ui[h] = ld_dev16(cur.imax, 4 * kPersistMaxW, 4 * (c0 + 4 * h));
                fa[h] = __builtin_bit_cast(floatx4, ld_dev16(cur.amprow, 4 * kPersistMaxW, 4 * (c0 + 4 * h)));
            }
            float m = -INFINITY;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const bool in = c0 + q < W;
                cv[q] = in ? fc[q >> 2][q & 3] : -INFINITY;
                iv[q] = (int)ui[q >> 2][q & 3];
                av[q] = fa[q >> 2][q & 3];
                cf[pad32(H + c0 + q)] = cv[q];
                if (in) {
                    rim[c0 + q] = iv[q];
                    ram[c0 + q] = av[q];
                }
                m = fmaxf(m, cv[q]);
            }
            bm[kBmHalo + tid] = m;
        }
        for (int j = tid; j < nown; j += kPT) cove[j] = covo[j] = -1;
        __syncthreads();
        const uint32_t bits = tid < nbk ? peak_bits8(cf, bm, tid, W, nt, a.th2) : 0u;
        const int cnt = __popc(bits);
        const int inc = wave_incl_scan(cnt);
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int woff = 0, n = 0;
#pragma unroll
        for (int k = 0; k < kPW; ++k) {
            const int vv = wsum[k];
            if (k < wave) woff += vv;
            n += vv;
        }
        if (n == 0) {
            if (lead) a.counters[3] = 1;
            break;
        }
        // the pass's list (ascending columns: list index = index in the pass) of the spikes applied,
        // and the owners' entries of the stored spike list
        const int nu = max(0, min(n, a.cap - base));   // spikes stored, as update_kernel applies them
        {
            int gi = woff + inc - cnt;   // peaks before this thread's block
            int r = r0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if ((bits >> q) & 1u) {
                    const int c = c0 + q;
                    const int ic = rim[c];
                    const float ac = ram[c];
                    if (gi < nu) {
                        lct[gi] = c | (ic << 16);
                        lamp[gi] = ac;
                        // own columns g + G j within [c - nt, c + nt]: the spike covers them
                        const int w0 = max(c - nt, 0), w1 = min(c + nt, W - 1);
                        const int jlo = w0 <= g ? 0 : floor_div(w0 - g + G - 1, G, rG);
                        const int jhi = min(w1 < g ? -1 : floor_div(w1 - g, G, rG), nown - 1);
                        for (int j = jlo; j <= jhi; ++j) atomicMax((gi & 1) ? covo + j : cove + j, gi);
                    }
                    if (r == g && base + gi < a.cap) {
                        a.sp_col[base + gi] = c;
                        a.sp_tmp[base + gi] = ic;
                        a.sp_amp[base + gi] = ac;
                    }
                    ++gi;
                }
                if (++r == G) r = 0;
            }
        }
        if (lead) {
            if (base + n > a.cap) a.counters[2] = 1;
            a.pass_off[np] = base;
            a.pass_off[np + 1] = min(base + n, a.cap);
        }
        ++np;
        __syncthreads();   // the list is complete
        PursuitArgs nxt = a;
        use_row(nxt, a, par ^ 1);
        // own columns, one thread each: carried into the next row when the pass does not rescore
        // them (uncovered, or covered within nt of an edge); the covered ones, with their last even-
        // and odd-indexed covering spikes, listed in ascending order and spread over the wavefronts,
        // NC at a time
        int ncv = 0;
        for (int j0 = 0; j0 < nown; j0 += kPT) {
            const int j = j0 + tid, wj = g + G * j;
            int e = -1, o = -1;
            if (j < nown) {
                e = cove[j];
                o = covo[j];
                if (!((e >= 0 || o >= 0) && wj >= nt && wj < W - nt)) {
                    st_dev(nxt.cfmax + wj, cf[pad32(H + wj)]);
                    st_dev(nxt.imax + wj, rim[wj]);
                    st_dev(nxt.amprow + wj, ram[wj]);
                }
            }
            const bool cvr = e >= 0 || o >= 0;
            const uint64_t m = __ballot(cvr);
            if (lane == 0) wsum[wave] = __popcll(m);
            __syncthreads();
            int off = ncv, tot = 0;
#pragma unroll
            for (int k = 0; k < kPW; ++k) {
                const int vv = wsum[k];
                if (k < wave) off += vv;
                tot += vv;
            }
            if (cvr) {
                const int k = off + __popcll(m & ((1ull << lane) - 1ull));
                cvj[k] = j;
                cvl[k] = e;
                cvp[k] = o;
            }
            ncv += tot;
            __syncthreads();   // the covered list (and wsum free again)
        }
        constexpr int NC = NV <= 3 ? 2 : 1;   // columns per step (registers)
        for (int k0 = NC * wave; k0 < ncv; k0 += NC * kPW) {
            int jj[NC], ww[NC], ee[NC], oo[NC], n2 = 0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                jj[c] = ww[c] = 0;
                ee[c] = oo[c] = -1;
                if (k0 + c < ncv) {
                    jj[c] = cvj[k0 + c];
                    ww[c] = g + G * jj[c];
                    ee[c] = cvl[k0 + c];
                    oo[c] = cvp[k0 + c];
                    ++n2;
                }
            }
            persist_update_pair<NV, NC>(a, nxt, psm4, ld4, lct, lamp, n2, jj, ww, ee, oo, nm4, rn4);
        }
        base += n;
        par ^= 1;
        cur = nxt;
        if (pass + 1 < max_passes && !grid_barrier(a, bbase, gen++, &s_ok, g, G)) return;
    }
    if (lead) {
        a.counters[0] = base;
        a.counters[1] = 0;
        a.counters[4] = np;
    }
    __syncthreads();   // every wavefront's rows final
    for (int j = wave; j < nown; j += kPW) {
        floatx4* dst = (floatx4*)(a.B + (int64_t)(g + G * j) * a.ldT);
        for (int i = lane; i < ld4; i += 64) dst[i] = psm4[j * ld4 + i];
    }
    if (a.ts) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) a.ts[2 * g + 1] = __builtin_amdgcn_s_memrealtime();
    }
}

template <int NV>
__global__ __launch_bounds__(kPT) void pursuit_persistent_kernel(PursuitArgs a, int first_pass, int max_passes) {
    persistent_body<NV>(a, first_pass, max_passes, (int)blockIdx.x, (int)gridDim.x);
}

// ------------------------------------------------------------------------------------------------
// residual, once after the last pass.  A workgroup owns 16 columns; it finds, per pass, the spikes
// whose +-nt/2 window reaches its columns (binary search in the pass's column-sorted list), lists
// them in pursuit order in LDS, and each wavefront walks that list for its columns applying, pass
// by pass, the last even then the last odd covering spike:
//   X[w][c] -= amp * waves[t][w - col + nt/2][c]
// which is the residual the reference's per-pass updates leave (template_matching.py:197).  When
// the passes' spikes reaching the block exceed the list (kResCap; short templates over many passes),
// the passes are taken in consecutive runs that fit, the columns updated run after run in pass
// order, so the result is the same.
// ------------------------------------------------------------------------------------------------
constexpr int kResCols = 16;
constexpr int kResCap = 1024;       // listed spikes per workgroup and run of passes
constexpr int kResMaxPasses = 256;
constexpr int kResBatch = 8;        // wave-row loads in flight per column
// threads per workgroup (its wavefronts share the 16 columns): 512 for short windows (two columns per
// wavefront; a 3122-sample window's residual is then latency bound on few workgroups), 256 for long
// ones (more workgroups of their own list building)
constexpr int kResThreadsShort = 512, kResThreadsLong = 256;

// NVC: float4 channel chunks per lane (NVC*256 >= C); NT: threads per workgroup; blk: the workgroup's
// 16 columns
template <int NVC, int NT>
__device__ __forceinline__ void residual_body(const PursuitArgs& a, int blk) {
    __shared__ int s_first[kResMaxPasses], s_cnt[kResMaxPasses], s_off[kResMaxPasses + 1];
    __shared__ int l_col[kResCap], l_tmp[kResCap], l_pp[kResCap];   // l_pp = pass * 2 + parity
    __shared__ float l_amp[kResCap];
    __shared__ int l_apply[NT / 64][2 * kResMaxPasses];              // per wavefront
    __shared__ int s_run[2];                                         // passes [s_run[0], s_run[1])
    if (a.counters[2]) return;                  // spike list overflowed: the call reports an error
    const int P = a.counters[4];                // passes that found spikes
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = a.nt / 2;
    const int64_t w0 = (int64_t)blk * kResCols;
    const int lo = (int)(w0 - h), hi = (int)(w0 + kResCols - 1 + h);
    for (int p = tid; p < P; p += NT) {
        const int b0 = a.pass_off[p], b1 = a.pass_off[p + 1];
        int l = b0, r = b1;                     // first index with col >= lo
        while (l < r) {
            const int m = (l + r) >> 1;
            if (a.sp_col[m] < lo) l = m + 1; else r = m;
        }
        int l2 = l, r2 = b1;                    // first index with col > hi
        while (l2 < r2) {
            const int m = (l2 + r2) >> 1;
            if (a.sp_col[m] <= hi) l2 = m + 1; else r2 = m;
        }
        s_first[p] = l;
        s_cnt[p] = l2 - l;
    }
    __syncthreads();
    constexpr int nvc = NVC;
    int* apply = l_apply[wave];
    // every thread keeps the run's end in a register: s_run is rewritten only after the barriers
    // that follow every thread's read of it
    for (int next = 0; next < P;) {
        if (tid == 0) {
            // the next run of passes whose spikes fit the list (at least one pass; a single pass
            // beyond the list, possible only with tied peaks, is cut and reported)
            const int p0 = next;
            int acc = 0, p = p0;
            while (p < P && (p == p0 || acc + s_cnt[p] <= kResCap)) {
                s_off[p] = acc;
                acc += s_cnt[p];
                ++p;
            }
            s_off[p] = acc;
            if (acc > kResCap) a.counters[5] = 1;
            s_run[0] = p0;
            s_run[1] = p;
        }
        __syncthreads();
        const int p0 = s_run[0], p1 = s_run[1];
        const int total = min(s_off[p1], kResCap);
        for (int p = p0 + tid; p < p1; p += NT) {
            for (int k = 0; k < s_cnt[p]; ++k) {
                const int d = s_off[p] + k;
                if (d >= kResCap) break;
                const int j = s_first[p] + k;
                l_col[d] = a.sp_col[j];
                l_tmp[d] = a.sp_tmp[j];
                l_amp[d] = a.sp_amp[j];
                l_pp[d] = p * 2 + ((j - a.pass_off[p]) & 1);
            }
        }
        __syncthreads();
        for (int cw = wave; cw < kResCols; cw += NT / 64) {
            const int64_t w = w0 + cw;
            if (w >= a.W) break;
            // 1) the ordered list of subtractions for this column: per pass, last even then last odd
            int na = 0, cur = -1, e = -1, o = -1;
            for (int d = 0; d <= total; ++d) {
                const bool end = d == total;
                const int pp = end ? -2 : l_pp[d];
                const bool cov = !end && (int)(w - l_col[d]) <= h && (int)(l_col[d] - w) <= h;
                if (end || (cov && (pp >> 1) != cur)) {
                    if (e >= 0) apply[na++] = e;
                    if (o >= 0) apply[na++] = o;
                    e = o = -1;
                    if (!end) cur = pp >> 1;
                }
                if (cov) {
                    if (pp & 1) o = d; else e = d;
                }
            }
            if (na == 0) continue;
            // 2) apply them in order, with the wave rows of 8 subtractions loaded ahead
            float* xr = a.X + w * a.ldC;
            floatx4 x[NVC];
#pragma unroll
            for (int v = 0; v < nvc; ++v) {
                const int c = 4 * lane + 256 * v;
                if (c < a.C) x[v] = *(const floatx4*)(xr + c);
            }
            for (int b0 = 0; b0 < na; b0 += kResBatch) {
                floatx4 wv[kResBatch][NVC];
#pragma unroll
                for (int q = 0; q < kResBatch; ++q) {
                    if (b0 + q < na) {
                        const int d = apply[b0 + q];
                        const float* wr = a.waves + ((int64_t)l_tmp[d] * a.nt + (w - l_col[d] + h)) * a.ldC;
#pragma unroll
                        for (int v = 0; v < nvc; ++v) {
                            const int c = 4 * lane + 256 * v;
                            if (c < a.C) wv[q][v] = *(const floatx4*)(wr + c);
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < kResBatch; ++q) {
                    if (b0 + q < na) {
                        const float am = l_amp[apply[b0 + q]];
#pragma unroll
                        for (int v = 0; v < nvc; ++v)
#pragma unroll
                            for (int r = 0; r < 4; ++r) x[v][r] = x[v][r] - am * wv[q][v][r];
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < nvc; ++v) {
                const int c = 4 * lane + 256 * v;
                if (c < a.C) *(floatx4*)(xr + c) = x[v];
            }
        }
        __syncthreads();   // the list is rewritten by the next run
        next = p1;
    }
}

template <int NVC, int NT>
__global__ __launch_bounds__(NT) void residual_kernel(PursuitArgs a) {
    residual_body<NVC, NT>(a, (int)blockIdx.x);
}

// ------------------------------------------------------------------------------------------------
hipError_t launch_score_all(const PursuitArgs& a, hipStream_t s) {
    const unsigned nb = (unsigned)std::min<int64_t>((a.W + 3) / 4, 4096);   // 4 columns per workgroup and round
    const int nv = (a.T + 255) / 256;
    switch (nv) {
#define LSS_SA(V) case V: hipLaunchKernelGGL(score_all_kernel<V>, dim3(nb), dim3(256), 0, s, a); break;
        LSS_SA(1) LSS_SA(2) LSS_SA(3) LSS_SA(4) LSS_SA(5) LSS_SA(6) LSS_SA(7) LSS_SA(8)
#undef LSS_SA
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_peaks(const PursuitArgs& a, hipStream_t s) {
    if (a.nt > kPeakHalo) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((a.W + kFlagCols - 1) / kFlagCols);
    hipLaunchKernelGGL(peaks_kernel, dim3(nb), dim3(kFlagCols), 0, s, a);
    return hipGetLastError();
}

// Measurement mode: per launch p, the first workgroup start and last workgroup end of the stamps
// ts[p][b][2] (b < grid) into out[2p], out[2p + 1]; one workgroup per launch.
__global__ __launch_bounds__(256) void stamp_span_kernel(const unsigned long long* __restrict__ ts, int grid,
                                                         unsigned long long* __restrict__ out) {
    __shared__ unsigned long long s0[256], s1[256];
    const unsigned long long* t = ts + (size_t)blockIdx.x * 2 * grid;
    unsigned long long a = ~0ull, b = 0;
    for (int i = threadIdx.x; i < grid; i += 256) {
        a = min(a, t[2 * i]);
        b = max(b, t[2 * i + 1]);
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            s0[threadIdx.x] = min(s0[threadIdx.x], s0[threadIdx.x + h]);
            s1[threadIdx.x] = max(s1[threadIdx.x], s1[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = s0[0];
        out[2 * blockIdx.x + 1] = s1[0];
    }
}

hipError_t launch_stamp_span(const unsigned long long* ts, int launches, int grid, unsigned long long* out,
                             hipStream_t s) {
    if (launches < 1 || grid < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(stamp_span_kernel, dim3(launches), dim3(256), 0, s, ts, grid, out);
    return hipGetLastError();
}

template <int NV, int CPW>
static void update_launch(const PursuitArgs& a, int grid, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
    // measurement mode: events recorded on the stream around the launch (their gap also holds the
    // dispatch latency; the workgroups' GPU-clock stamps give the execution time)
    if (start) (void)hipEventRecord(start, s);
    hipLaunchKernelGGL((update_kernel<NV, CPW>), dim3(grid), dim3(256), 0, s, a);
    if (stop) (void)hipEventRecord(stop, s);
}

hipError_t launch_update(const PursuitArgs& a, int grid, hipStream_t s, hipEvent_t start, hipEvent_t stop, int cpw) {
    const int nv = (a.T + 255) / 256;
    if (grid < 1) grid = kUpdateGrid;
    if (cpw != 1 && cpw != 2) return hipErrorInvalidValue;
    if (nv > 4) cpw = 1;   // more than 1024 templates: one score column per wavefront (registers)
    switch (nv * 2 + cpw - 1) {
#define LSS_UP(V)                                                            \
    case 2 * V: update_launch<V, 1>(a, grid, s, start, stop); break;         \
    case 2 * V + 1: update_launch<V, 2>(a, grid, s, start, stop); break;
        LSS_UP(1) LSS_UP(2) LSS_UP(3) LSS_UP(4)
#undef LSS_UP
        case 10: update_launch<5, 1>(a, grid, s, start, stop); break;
        case 12: update_launch<6, 1>(a, grid, s, start, stop); break;
        case 14: update_launch<7, 1>(a, grid, s, start, stop); break;
        case 16: update_launch<8, 1>(a, grid, s, start, stop); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// LDS of the persistent pursuit at grid G (bytes) and the most a workgroup may take (gfx950: 160 KB,
// less the kernel's static arrays)
constexpr size_t kPersistLdsMax = 160 * 1024 - 256;
size_t persistent_lds(int64_t W, int64_t ldT, int G) {
    return sizeof(float) * (size_t)PersistLayout((int)W, (int)ldT, G).words;
}

typedef void (*PersistentFn)(PursuitArgs, int, int);
PersistentFn persistent_fn(int T) {
    switch ((T + 255) / 256) {
        case 1: return pursuit_persistent_kernel<1>;
        case 2: return pursuit_persistent_kernel<2>;
        case 3: return pursuit_persistent_kernel<3>;
        case 4: return pursuit_persistent_kernel<4>;
        default: return nullptr;
    }
}

bool persistent_supports(int T) { return persistent_fn(T) != nullptr; }

// workgroups of fn with `lds` bytes each that the device holds at once
static hipError_t resident_blocks(PersistentFn fn, size_t lds, int* n) {
    int dev = 0, cus = 0, per = 0;
    *n = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPersistLdsMax);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)fn, kPT, lds);
    if (e == hipSuccess) *n = per * cus;
    return e;
}

// The largest grid g <= want whose workgroups, with the LDS that grid needs (fewer workgroups hold
// more score rows each), the device holds at once; 0 if none does.
static hipError_t coresident_grid(PersistentFn fn, int64_t W, int64_t ldT, int want, int* grid) {
    *grid = 0;
    int g = want;
    for (int it = 0; it < 16 && g >= 1; ++it) {
        const size_t lds = persistent_lds(W, ldT, g);
        if (lds > kPersistLdsMax) return hipSuccess;   // a smaller grid needs more LDS still
        int cap = 0;
        const hipError_t e = resident_blocks(fn, lds, &cap);
        if (e != hipSuccess) return e;
        if (cap >= g) {
            *grid = g;
            return hipSuccess;
        }
        g = cap;
    }
    return hipSuccess;
}

hipError_t persistent_max_grid(int T, int64_t W, int* grid) {
    const PersistentFn fn = persistent_fn(T);
    *grid = 0;
    if (!fn) return hipSuccess;   // more than 1024 templates: per-pass kernels only
    if (W < 1 || W > kPersistMaxW) return hipErrorInvalidValue;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    return coresident_grid(fn, W, (T + 3) / 4 * 4, 4 * cus, grid);
}

hipError_t persistent_grid(const PursuitArgs& a, int want, int* grid) {
    const PersistentFn fn = persistent_fn(a.T);
    *grid = 0;
    if (!fn || a.W < 1 || a.W > kPersistMaxW || want < 1) return hipSuccess;
    return coresident_grid(fn, a.W, a.ldT, want, grid);
}

hipError_t launch_pursuit_persistent(const PursuitArgs& a, int grid, int first_pass, int max_passes, hipStream_t s,
                                     int* used) {
    const PersistentFn fn = persistent_fn(a.T);
    if (!fn || a.W > kPersistMaxW || a.nt > kPeakHalo || grid < 1 || first_pass < 0 || first_pass >= max_passes ||
        !a.prow_cf || !a.prow_im || !a.prow_amp || !a.bar || a.ldT % 4 != 0)
        return hipErrorInvalidValue;
    int g = 0;
    const hipError_t e = coresident_grid(fn, a.W, a.ldT, std::min(grid, kPersistMaxGrid), &g);
    if (e != hipSuccess) return e;
    if (g < 1) return hipErrorInvalidConfiguration;   // the score rows do not fit the LDS of that grid
    if (used) *used = g;
    hipLaunchKernelGGL(fn, dim3(g), dim3(kPT), persistent_lds(a.W, a.ldT, g), s, a, first_pass, max_passes);
    return hipGetLastError();
}

hipError_t launch_residual(const PursuitArgs& a, hipStream_t s) {
    if (a.C > 1024) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((a.W + kResCols - 1) / kResCols);
    const bool shrt = a.W <= kPersistMaxW;
    const unsigned nt = shrt ? kResThreadsShort : kResThreadsLong;
#define LSS_RES(V)                                                                          \
    if (shrt)                                                                               \
        hipLaunchKernelGGL((residual_kernel<V, kResThreadsShort>), dim3(nb), dim3(nt), 0, s, a); \
    else                                                                                    \
        hipLaunchKernelGGL((residual_kernel<V, kResThreadsLong>), dim3(nb), dim3(nt), 0, s, a)
    switch ((a.C + 255) / 256) {
        case 1: LSS_RES(1); break;
        case 2: LSS_RES(2); break;
        default: LSS_RES(4); break;
    }
#undef LSS_RES
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// PC features and positions (template_matching.py:50-51, postprocessing.py:33-46): one wavefront
// per spike.  The spike's nt x nC residual window is gathered into LDS with all loads in flight at
// once, then lane (j, k) forms sum_m X[w - nt/2 + m][iCC[j, iU[t]]] wPCA[k][m] + amp Ucc[j][t][k].
// ------------------------------------------------------------------------------------------------
// Workgroup g of G (grid-stride over the spikes).  stg (may be null): readback staging that the
// spikes below its capacity (column, template, amplitude, features, position) and the counters are
// also written to, as the staging kernel would write them.
__device__ __forceinline__ void features_body(const FeatureArgs& a, int g, int G, const StageArgs* stg = nullptr) {
    // wpca [K*nt] | per wavefront: window [nt][nC], features [nC*K], channel mass, x, y [nC]
    extern __shared__ float fsm[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int n = a.counters ? a.counters[0] + a.counters[1] : a.n;
    if (n > a.cap) n = a.cap;
    const int h = a.nt / 2, K = a.K, nC = a.nC, nt = a.nt;
    float* wp = fsm;
    float* xwin = fsm + K * nt + wave * (nt * nC + nC * K + 3 * nC);
    float* sf = xwin + nt * nC;
    float* sm = sf + nC * K;
    float* sx = sm + nC;
    float* sy = sx + nC;
    for (int e = threadIdx.x; e < K * nt; e += 256) wp[e] = a.wpca[e];
    if (stg && g == 0 && threadIdx.x < 8) stg->h_counters[threadIdx.x] = a.counters[threadIdx.x];
    __syncthreads();
    for (int s0 = g * 4; s0 < n; s0 += G * 4) {
        const int s = s0 + wave;
        if (s < n) {
            const int w = a.sp_col[s];
            const int t = a.sp_tmp[s];
            const float am = a.sp_amp[s];
            const int iu = a.iU[t];
            const bool staged = stg && s < stg->cap;
            if (staged && lane == 0) {
                stg->h_col[s] = w;
                stg->h_tmp[s] = t;
                stg->h_amp[s] = am;
            }
            // lane (r, j): channel j of rows r, r + rpi, ... (rpi = 64 / nC rows per round); its
            // channel index is loaded once, every row load is independent of the others
            {
                const int rpi = 64 / nC, r = lane / nC, j = lane - r * nC;
                if (r < rpi) {
                    const int ch = a.iCC[j * a.C + iu];
                    for (int m = r; m < nt; m += rpi) {
                        const int64_t ww = (int64_t)w - h + m;
                        xwin[m * nC + j] = (ww >= 0 && ww < a.W) ? a.X[ww * a.ldC + ch] : 0.f;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int p = lane; p < nC * K; p += 64) {
                const int j = p / K, k = p % K;
                float acc = 0.f;
                for (int m = 0; m < nt; ++m) acc = acc + xwin[m * nC + j] * wp[k * nt + m];
                const float f = acc + am * a.Ucc[((int64_t)j * a.T + t) * K + k];
                sf[p] = f;
                a.tF[(int64_t)s * nC * K + p] = f;
                if (staged) stg->h_tF[(int64_t)s * nC * K + p] = f;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // lane j: channel j's mass sum_k F^2 and site; lane 0 then sums in channel order
            if (lane < nC) {
                float m = 0.f;
                for (int k = 0; k < K; ++k) m += sf[lane * K + k] * sf[lane * K + k];
                const int ch = a.iCC[lane * a.C + iu];
                sm[lane] = m;
                sx[lane] = a.xc[ch];
                sy[lane] = a.yc[ch];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0) {
                float tot = 0.f;
                for (int j = 0; j < nC; ++j) tot += sm[j];
                float xs = 0.f, ys = 0.f;
                for (int j = 0; j < nC; ++j) {
                    const float wgt = sm[j] / tot;
                    xs += sx[j] * wgt;
                    ys += sy[j] * wgt;
                }
                a.xs[s] = xs;
                a.ys[s] = ys;
                if (staged) {
                    stg->h_xs[s] = xs;
                    stg->h_ys[s] = ys;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

__global__ __launch_bounds__(256) void features_kernel(FeatureArgs a) {
    features_body(a, (int)blockIdx.x, (int)gridDim.x);
}

hipError_t launch_features(const FeatureArgs& a, int grid, hipStream_t s) {
    if (a.nC * a.K > 32 * 8 || a.nC > 32) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * ((size_t)a.K * a.nt + 4 * ((size_t)a.nt * a.nC + a.nC * a.K + 3 * a.nC));
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)features_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(features_kernel, dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Readback staging: the counters and the first min(n, cap) spikes (n = counters[0] + counters[1])
// written straight into pinned host buffers, so that a replayed graph moves only the spikes found.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void stage_body(const StageArgs& a, int g, int G) {
    int n = a.counters[0] + a.counters[1];
    n = min(n, a.cap);
    const int gid = g * 256 + threadIdx.x, stride = G * 256;
    if (gid < 8) a.h_counters[gid] = a.counters[gid];
    for (int i = gid; i < n; i += stride) {
        a.h_col[i] = a.col[i];
        a.h_tmp[i] = a.tmp[i];
        a.h_amp[i] = a.amp[i];
        a.h_xs[i] = a.xs[i];
        a.h_ys[i] = a.ys[i];
    }
    const int64_t nf = (int64_t)n * a.FK;
    for (int64_t e = gid; e < nf; e += stride) a.h_tF[e] = a.tF[e];
}

__global__ __launch_bounds__(256) void stage_spikes_kernel(StageArgs a) {
    stage_body(a, (int)blockIdx.x, (int)gridDim.x);
}

hipError_t launch_stage_spikes(const StageArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(stage_spikes_kernel, dim3(64), dim3(256), 0, s, a);
    return hipGetLastError();
}

// counters[0..7] = 0 (a fill kernel, so that the sequence can be captured in a graph)
// Counters zeroed and the persistent pursuit's barrier epoch advanced; every 2^16 epochs the barrier
// flags are raised to the new epoch's base, so that no flag falls 2^31 behind a barrier's target.
__device__ __forceinline__ void reset_body(int* c, int* bar) {
    const int e = c[kEpoch] + 1;
    __syncthreads();
    if (threadIdx.x < kCounters) c[threadIdx.x] = threadIdx.x == kEpoch ? e : 0;
    if (bar && (e & 0xFFFF) == 0)
        for (int i = threadIdx.x; i < kPersistMaxGrid; i += blockDim.x) bar[kBarStride * i] = (int)((unsigned)e * 4096u);
}

__global__ void reset_counters_kernel(int* c, int* bar) { reset_body(c, bar); }

hipError_t launch_reset_counters(int* counters, hipStream_t s, int* bar) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counters, bar);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Grouped launches: the windows of several handles (members) in one launch per stage, member i on
// workgroups [first[i], first[i + 1]), each running the single-window kernel's body on its share.
// The member descriptors are read from device memory; the span (kernel argument) is indexed only
// with constants.
// ------------------------------------------------------------------------------------------------
struct GroupSpan {
    int n;
    int first[kGroupMax + 1];
    int aux[kGroupMax];   // persistent pursuit: the member's max_passes
};

// The member this workgroup belongs to, its first workgroup g0, its end g1 and its aux entry.
__device__ __forceinline__ int group_member(const GroupSpan& sp, int* g0, int* g1, int* aux) {
    int m = 0, a0 = 0, a1 = sp.first[1], x = sp.aux[0];
#pragma unroll
    for (int i = 1; i < kGroupMax; ++i)
        if (i < sp.n && (int)blockIdx.x >= sp.first[i]) {
            m = i;
            a0 = sp.first[i];
            a1 = sp.first[i + 1];
            x = sp.aux[i];
        }
    *g0 = a0;
    *g1 = a1;
    *aux = x;
    return m;
}

__global__ void reset_group_kernel(const GroupMember* __restrict__ d) {
    reset_body(d[blockIdx.x].a.counters, d[blockIdx.x].a.bar);
}

__global__ __launch_bounds__(kPT) void pursuit_persistent_group_kernel(const GroupMember* __restrict__ d, GroupSpan sp) {
    int g0, g1, mp;
    const int m = group_member(sp, &g0, &g1, &mp);
    const PursuitArgs a = d[m].a;
    const int g = (int)blockIdx.x - g0, G = g1 - g0;
    switch ((a.T + 255) >> 8) {
        case 1: persistent_body<1>(a, 0, mp, g, G); break;
        case 2: persistent_body<2>(a, 0, mp, g, G); break;
        case 3: persistent_body<3>(a, 0, mp, g, G); break;
        case 4: persistent_body<4>(a, 0, mp, g, G); break;
        default: break;   // rejected by the launcher
    }
}

template <int NVC>   // for the widest member: NVC*256 >= C of every member (short windows)
__global__ __launch_bounds__(kResThreadsShort) void residual_group_kernel(const GroupMember* __restrict__ d,
                                                                          GroupSpan sp) {
    int g0, g1, x;
    const int m = group_member(sp, &g0, &g1, &x);
    const PursuitArgs a = d[m].a;
    residual_body<NVC, kResThreadsShort>(a, (int)blockIdx.x - g0);
}

__global__ __launch_bounds__(256) void features_group_kernel(const GroupMember* __restrict__ d, GroupSpan sp) {
    int g0, g1, x;
    const int m = group_member(sp, &g0, &g1, &x);
    const FeatureArgs f = d[m].f;
    const StageArgs stg = d[m].s;
    features_body(f, (int)blockIdx.x - g0, g1 - g0, &stg);
}

static GroupSpan even_span(int n, int each) {
    GroupSpan sp{};
    sp.n = n;
    for (int i = 0; i <= n; ++i) sp.first[i] = i * each;
    return sp;
}

hipError_t launch_group_reset(const GroupMember* d, int n, hipStream_t s) {
    if (!d || n < 1 || n > kGroupMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reset_group_kernel, dim3(n), dim3(64), 0, s, d);
    return hipGetLastError();
}

// LDS per workgroup of a grouped persistent launch (the largest member's) and its total grid; false
// when a member cannot run in it.
static bool group_persistent_shape(const GroupMember* m, const int* grids, int n, size_t* lds, int* total) {
    *lds = 0;
    *total = 0;
    if (!m || !grids || n < 1 || n > kGroupMax) return false;
    for (int i = 0; i < n; ++i) {
        const PursuitArgs& a = m[i].a;
        if (!persistent_fn(a.T) || a.W < 1 || a.W > kPersistMaxW || a.nt > kPeakHalo || grids[i] < 1 ||
            grids[i] > kPersistMaxGrid || !a.prow_cf || !a.prow_im || !a.prow_amp || !a.bar || a.ldT % 4 != 0)
            return false;
        *lds = std::max(*lds, persistent_lds(a.W, a.ldT, grids[i]));
        *total += grids[i];
    }
    return *lds <= kPersistLdsMax;
}

hipError_t group_persistent_fits(const GroupMember* m, const int* grids, int n, bool* fits) {
    *fits = false;
    size_t lds = 0;
    int total = 0;
    if (!group_persistent_shape(m, grids, n, &lds, &total)) return hipSuccess;
    int dev = 0, cus = 0, per = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)pursuit_persistent_group_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPersistLdsMax);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)pursuit_persistent_group_kernel, kPT, lds);
    if (e == hipSuccess) *fits = per * cus >= total;
    return e;
}

hipError_t launch_group_persistent(const GroupMember* d, const GroupMember* m, const int* grids, const int* max_passes,
                                   int n, hipStream_t s) {
    size_t lds = 0;
    int total = 0;
    if (!d || !max_passes || !group_persistent_shape(m, grids, n, &lds, &total)) return hipErrorInvalidValue;
    bool fits = false;
    const hipError_t e = group_persistent_fits(m, grids, n, &fits);
    if (e != hipSuccess) return e;
    if (!fits) return hipErrorInvalidConfiguration;   // not every workgroup would be resident at once
    GroupSpan sp{};
    sp.n = n;
    for (int i = 0; i < n; ++i) {
        if (max_passes[i] < 1) return hipErrorInvalidValue;
        sp.first[i + 1] = sp.first[i] + grids[i];
        sp.aux[i] = max_passes[i];
    }
    hipLaunchKernelGGL(pursuit_persistent_group_kernel, dim3(total), dim3(kPT), lds, s, d, sp);
    return hipGetLastError();
}

hipError_t launch_group_residual(const GroupMember* d, const GroupMember* m, int n, hipStream_t s) {
    if (!d || !m || n < 1 || n > kGroupMax) return hipErrorInvalidValue;
    GroupSpan sp{};
    sp.n = n;
    int cmax = 0;
    for (int i = 0; i < n; ++i) {
        if (m[i].a.C > 1024 || m[i].a.W < 1) return hipErrorInvalidValue;
        cmax = std::max(cmax, m[i].a.C);
        sp.first[i + 1] = sp.first[i] + (int)((m[i].a.W + kResCols - 1) / kResCols);
    }
    const dim3 grid(sp.first[n]);
    switch ((cmax + 255) / 256) {
        case 1: hipLaunchKernelGGL(residual_group_kernel<1>, grid, dim3(kResThreadsShort), 0, s, d, sp); break;
        case 2: hipLaunchKernelGGL(residual_group_kernel<2>, grid, dim3(kResThreadsShort), 0, s, d, sp); break;
        default: hipLaunchKernelGGL(residual_group_kernel<4>, grid, dim3(kResThreadsShort), 0, s, d, sp); break;
    }
    return hipGetLastError();
}

hipError_t launch_group_features(const GroupMember* d, const GroupMember* m, int n, hipStream_t s) {
    if (!d || !m || n < 1 || n > kGroupMax) return hipErrorInvalidValue;
    size_t lds = 0;
    for (int i = 0; i < n; ++i) {
        const FeatureArgs& a = m[i].f;
        if (a.nC * a.K > 32 * 8 || a.nC > 32 || !a.counters) return hipErrorInvalidValue;
        lds = std::max(lds, sizeof(float) * ((size_t)a.K * a.nt + 4 * ((size_t)a.nt * a.nC + a.nC * a.K + 3 * a.nC)));
    }
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute((const void*)features_group_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(features_group_kernel, dim3(n * kGroupFeatureGrid), dim3(256), lds, s, d,
                       even_span(n, kGroupFeatureGrid));
    return hipGetLastError();
}

}  // namespace lss
