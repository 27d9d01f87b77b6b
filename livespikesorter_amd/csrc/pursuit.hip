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
// every template a candidate.  With DEV, b and m receive the winner's B[w][t] and nm_t.
constexpr float kCandidate = 0.99999905f;   // 1 - 2^-20
template <int NV, bool DEV>
__device__ __forceinline__ void rescore(const floatx4* x, const floatx4* nmv, const floatx4* rnv, int T, float& best,
                                        int& bi, float& bb, float& bm) {
    const int lane = threadIdx.x & 63;
    float sa[NV][4];
    float ml = -1.f;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float r = fmaxf(x[i][e], 0.f);
            const float sv = (r * r) * rnv[i][e];
            sa[i][e] = (4 * lane + 256 * i + e < T) ? sv : -1.f;
            ml = fmaxf(ml, sa[i][e]);
        }
    const float M = wave_max(ml);
    const float thr = (M >= 1e-30f && M <= 1e30f) ? M * kCandidate : -0.5f;
    best = -1.f;
    bi = 0x7fffffff;
    bb = 0.f;
    bm = 1.f;
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

// Device-scope (all XCDs) relaxed loads and stores: the only way workgroups of one launch of the
// persistent pursuit exchange data, so that its grid barrier needs no cache write-back/invalidate.
template <typename V>
__device__ __forceinline__ V ld_dev(const V* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename V>
__device__ __forceinline__ void st_dev(V* p, V v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Score rows B[w][t .. t+3]: plain 16-byte accesses between kernels; inside the persistent pursuit
// (DEV), whose workgroups may update any column in any pass, device-coherent ones (the same cache
// policy as the device-scope loads and stores above) through a buffer resource over B, whose bounds
// the hardware checks (offsets past the end read 0 and drop the store).
constexpr int kCoherent = 16;          // buffer cache policy: device scope
constexpr int kRawBufferWord3 = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t b_rsrc(const PursuitArgs& a) {
    const int64_t bytes = a.W * a.ldT * 4;
    return __builtin_amdgcn_make_buffer_rsrc(a.B, (short)0, (int)min(bytes, (int64_t)INT_MAX), kRawBufferWord3);
}
template <bool DEV>
__device__ __forceinline__ floatx4 ld_b(const PursuitArgs& a, __amdgpu_buffer_rsrc_t r, int64_t w, int t) {
    if constexpr (DEV)
        return __builtin_bit_cast(floatx4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)((w * a.ldT + t) * 4), 0,
                                                                                 kCoherent));
    else return *(const floatx4*)(a.B + w * a.ldT + t);
}
template <bool DEV>
__device__ __forceinline__ void st_b(const PursuitArgs& a, __amdgpu_buffer_rsrc_t r, int64_t w, int t, floatx4 v) {
    if constexpr (DEV)
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(uint32x4, v), r, (int)((w * a.ldT + t) * 4), 0,
                                               kCoherent);
    else *(floatx4*)(a.B + w * a.ldT + t) = v;
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
template <int NV, bool DEV = false>   // float4 vectors per lane: NV*256 >= T
__device__ __forceinline__ void score_column(const PursuitArgs& a, int64_t w, int lane, const floatx4* nmv,
                                             const floatx4* rnv) {
    float best = 0.f, bb = 0.f, bm = 1.f;
    int bi = 0;
    if (w >= a.nt && w < a.W - a.nt) {
        const __amdgpu_buffer_rsrc_t r = b_rsrc(a);
        floatx4 b[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int t = 4 * lane + 256 * i;
            b[i] = (t < a.T) ? ld_b<DEV>(a, r, w, t) : floatx4{0.f, 0.f, 0.f, 0.f};
        }
        rescore<NV, DEV>(b, nmv, rnv, a.T, best, bi, bb, bm);
    }
    if (lane == 0) store_score<DEV>(a, w, best, bi, bb, bm);
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
constexpr int kColsB = 8;     // column alignment of the persistent pursuit's ownership runs
constexpr int kUpdateGrid = 8192;   // workgroups per update launch (grid-stride over work items)

// A pass's spike list (ascending columns): plain loads (a list stored by an earlier kernel, or in
// LDS) or device-scope loads (a list another workgroup of the same launch stored).
struct ListPlain {
    const int* col;
    const int* tmp;
    const float* amp;
    __device__ int c(int i) const { return col[i]; }
    __device__ int t(int i) const { return tmp[i]; }
    __device__ float a(int i) const { return amp[i]; }
};
struct ListDev {
    const int* col;
    const int* tmp;
    const float* amp;
    __device__ int c(int i) const { return ld_dev(col + i); }
    __device__ int t(int i) const { return ld_dev(tmp + i); }
    __device__ float a(int i) const { return ld_dev(amp + i); }
};

// Spikes of the pass covering column w within +-reach, starting at spike s (whose window covers
// w): the last even-indexed (e) and odd-indexed (o) ones.  Spikes are in ascending column order,
// so the covering ones are s, s+1, ...; c1, c2 are col[s+1], col[s+2] (INT_MAX past the end).
template <typename L>
__device__ __forceinline__ void covering(const L& l, int s, int n, int c1, int c2, int w, int reach,
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
template <int NV, typename List, bool DEV>   // float4 vectors per lane over templates: NV*256 >= T
__device__ __forceinline__ void update_two(const PursuitArgs& a, const List& l, int n, int wa, int sa, int ca1,
                                           int ca2, bool va, int wb, int sb, int cb1, int cb2, bool vb,
                                           const floatx4* nmv, const floatx4* rnv) {
    const int lane = threadIdx.x & 63;
    const int nt = a.nt, L = 2 * nt + 1;
    // two columns per wavefront, loads of both issued before any use
    int ea = -1, oa = -1, eb = -1, ob = -1;
    if (va) covering(l, sa, n, ca1, ca2, wa, nt, ea, oa);
    if (vb) covering(l, sb, n, cb1, cb2, wb, nt, eb, ob);
    floatx4 ba[NV], bb[NV], cea[NV], coa[NV], ceb[NV], cob[NV];
    const __amdgpu_buffer_rsrc_t r = b_rsrc(a);
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
                ba[i] = ld_b<DEV>(a, r, wa, t);
                if (qea) cea[i] = *(const floatx4*)(qea + t);
                if (qoa) coa[i] = *(const floatx4*)(qoa + t);
            }
            if (vb) {
                bb[i] = ld_b<DEV>(a, r, wb, t);
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
            if (t < a.T) st_b<DEV>(a, r, wa, t, ba[i]);
        }
        if (vb) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float v = bb[i][e];
                if (qeb) v = v - aeb * ceb[i][e];
                if (qob) v = v - aob * cob[i][e];
                bb[i][e] = v;
            }
            if (t < a.T) st_b<DEV>(a, r, wb, t, bb[i]);
        }
    }
    // rescored inside [nt, W - nt) only (the reference's Cf there; the edges never hold a peak)
    if (va && wa >= nt && wa < a.W - nt) {
        float best, b_, m_;
        int bi;
        rescore<NV, DEV>(ba, nmv, rnv, a.T, best, bi, b_, m_);
        if (lane == 0) store_score<DEV>(a, wa, best, bi, b_, m_);
    }
    if (vb && wb >= nt && wb < a.W - nt) {
        float best, b_, m_;
        int bi;
        rescore<NV, DEV>(bb, nmv, rnv, a.T, best, bi, b_, m_);
        if (lane == 0) store_score<DEV>(a, wb, best, bi, b_, m_);
    }
}

// Per-pass update: the chunk [c0, c0 + 4 CPW) of the columns [lo, w1] whose first covering spike is
// s (c1, c2 = the next two spikes' columns), CPW columns per wavefront.
template <int NV, typename List, bool DEV, int CPW>
__device__ __forceinline__ void update_score_columns(const PursuitArgs& a, const List& l, int s, int n, int c1, int c2,
                                                     int c0, int lo, int w1, const floatx4* nmv,
                                                     const floatx4* rnv) {
    const int wa = c0 + CPW * (int)(threadIdx.x >> 6), wb = wa + 1;
    update_two<NV, List, DEV>(a, l, n, wa, s, c1, c2, wa >= lo && wa <= w1, wb, s, c1, c2,
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
        if (max(c0, lo) <= w1) update_score_columns<NV, ListPlain, false, CPW>(a, l, s, n, c1, c2, c0, lo, w1, nmv, rnv);
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
// per pass.
//
// Workgroup g owns the columns [a0, a1) (a fixed partition): it alone scores them, reads their B
// rows and updates them, so B stays in the caches of one XCD and needs no coherence traffic.  What
// workgroups exchange goes through device-scope loads and stores only (no cache write-back or
// invalidate at the barrier): per column its Cfmax, imax and the amplitude B[w][imax]/nm[imax] it
// would have as a peak, in one of two rows chosen by pass parity.
//
// Pass p: every workgroup loads row p&1 (Cfmax, imax, amplitude of every column) into LDS, forms
// the exact window max by doubling (window_max_pow2), applies the reference's test and compacts
// the peaks in ascending order, so each holds the pass's list (columns, templates, amplitudes)
// without exchanging it; workgroup 0 stores the list.  Each then updates its own columns' B rows
// and publishes the rescored columns into row (p+1)&1, copying the values of its columns the pass
// leaves unchanged; the barrier ends the pass.  A workgroup writes row (p+1)&1 while others may
// still read row p&1, never the row they read, so the one barrier per pass orders every read of a
// pass before the writes that overwrite it.  Every workgroup sees the same row, so all take the
// same branches and leave the loop on the same pass.
//
// The barrier waits a bounded time: on a timeout (workgroups not co-resident) the kernel sets
// counters[7], every workgroup leaves, and the host reports the error; the launcher caps G at the
// occupancy limit so that this does not happen in normal operation.
// ------------------------------------------------------------------------------------------------
constexpr int kPT = 512;                                // threads per persistent workgroup
constexpr int kPW = kPT / 64;                           // its wavefronts
// LDS index of row entry i with one pad word per 32 (stride-8 reads conflict-free); pad_len(n) covers [0, n)
__device__ __forceinline__ int pad32(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int pad_len(int n) { return n + (n >> 5) + 1; }
constexpr int kBarrierSpin = 1 << 20;                   // ~1 s of polling
constexpr int kPersistLoads = (kPersistMaxW + 2 * kPeakHalo + kPT - 1) / kPT;   // row values per thread

// Grid barrier number `gen` (1, 2, ...) over counters[6] (zero at launch).  Data shared between
// workgroups is only accessed with device-scope atomics, so waiting for this wavefront's memory
// operations before the workgroup meets orders them before the arrival; no cache maintenance.
// Returns false after a timeout or when another workgroup has timed out.
__device__ __forceinline__ bool grid_barrier(int* counters, int gen, int* s_ok) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(counters + 6, 1);
        const int target = gen * (int)gridDim.x;
        int ok = 1;
        for (int it = 0; ld_dev(counters + 6) < target; ++it) {
            if (it >= kBarrierSpin || ld_dev(counters + 7)) {
                atomicExch(counters + 7, 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *s_ok = ok;
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

// First spike of an ascending list whose +-nt window reaches column w (n if none does).
__device__ __forceinline__ int first_cover(const ListPlain& l, int n, int w, int nt) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (l.col[mid] + nt < w) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The persistent pursuit's update, spread evenly over every wavefront of the launch.  The columns
// the pass changes are the union of its spikes' windows [col - nt, col + nt] within [0, W);
// spike s adds the run [st_s, en_s], st_s = max(col_s - nt, col_{s-1} + nt + 1, 0),
// en_s = min(col_s + nt, W - 1), and pfx[s] (exclusive prefix of the run lengths, pfx[n] = K) numbers
// them in ascending order.  Covered column k goes to wavefront k / 2 mod (G * kPW) of the launch, two
// per wavefront and round; its first covering spike is the s whose run holds it (no earlier
// spike's window reaches past en_{s-1}).  Each column gets the update update_items gives it and,
// inside [nt, W - nt), is rescored into the next row.
__device__ __forceinline__ int run_start(const ListPlain& l, int s, int nt) {
    return max(max(l.col[s] - nt, s > 0 ? l.col[s - 1] + nt + 1 : 0), 0);
}

// pfx[0 .. n] for the first n spikes of the list (n <= kPT * kPersistLoads); every thread returns K.
__device__ __forceinline__ int covered_prefix(const ListPlain& l, int n, int nt, int W, int* pfx, int* wsum) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (n + kPT - 1) / kPT;
    const int s0 = tid * per;
    int loc = 0;
    for (int q = 0; q < per; ++q) {
        const int sp = s0 + q;
        if (sp < n) loc += max(0, min(l.col[sp] + nt, W - 1) - run_start(l, sp, nt) + 1);
    }
    int inc = loc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(inc, off, 64);
        if (lane >= off) inc += v;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int woff = 0, K = 0;
#pragma unroll
    for (int k = 0; k < kPW; ++k) {
        const int v = wsum[k];
        if (k < wave) woff += v;
        K += v;
    }
    int run = woff + inc - loc;
    for (int q = 0; q < per; ++q) {
        const int sp = s0 + q;
        if (sp < n) {
            pfx[sp] = run;
            run += max(0, min(l.col[sp] + nt, W - 1) - run_start(l, sp, nt) + 1);
        }
    }
    if (tid == 0) pfx[n] = K;
    __syncthreads();
    return K;
}

// the covered column numbered k: its column and first covering spike (pfx[s] <= k < pfx[s + 1])
__device__ __forceinline__ int covered_column(const ListPlain& l, const int* pfx, int n, int nt, int k, int* s_out) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {   // last s with pfx[s] <= k
        const int mid = (lo + hi + 1) >> 1;
        if (pfx[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    *s_out = lo;
    return run_start(l, lo, nt) + (k - pfx[lo]);
}

template <int NV>
__device__ __forceinline__ void update_covered(const PursuitArgs& a, const ListPlain& l, int n, const int* pfx, int K,
                                               const floatx4* nmv, const floatx4* rnv) {
    const int nt = a.nt;
    const int gw = (int)blockIdx.x * kPW + (int)(threadIdx.x >> 6), nw = (int)gridDim.x * kPW;
    for (int k0 = 2 * gw; k0 < K; k0 += 2 * nw) {
        int sa = 0, sb = 0;
        const int wa = covered_column(l, pfx, n, nt, k0, &sa);
        const bool vb = k0 + 1 < K;
        const int wb = vb ? covered_column(l, pfx, n, nt, k0 + 1, &sb) : wa;
        const int ca1 = (sa + 1 < n) ? l.col[sa + 1] : INT_MAX / 2, ca2 = (sa + 2 < n) ? l.col[sa + 2] : INT_MAX / 2;
        const int cb1 = (sb + 1 < n) ? l.col[sb + 1] : INT_MAX / 2, cb2 = (sb + 2 < n) ? l.col[sb + 2] : INT_MAX / 2;
        update_two<NV, ListPlain, true>(a, l, n, wa, sa, ca1, ca2, true, wb, sb, cb1, cb2, vb, nmv, rnv);
    }
}

// The columns [a0, a1) the pass leaves unrescored (uncovered, or covered within nt of an edge) carry
// their values (cf, im, am: the pass's row in LDS; cf padded, with its halo) into the next row.
__device__ __forceinline__ void carry_own_columns(const PursuitArgs& a, const ListPlain& l, int n, int a0, int a1,
                                                  const float* cf, const int* im, const float* am) {
    const int nt = a.nt, W = (int)a.W;
    for (int w = a0 + (int)threadIdx.x; w < a1; w += kPT) {
        const int s = first_cover(l, n, w, nt);
        const bool cov = s < n && l.col[s] - nt <= w;
        if (!(cov && w >= nt && w < W - nt)) {
            st_dev(a.cfmax + w, cf[pad32(kPeakHalo + w)]);
            st_dev(a.imax + w, im[w]);
            st_dev(a.amprow + w, am[w]);
        }
    }
}

// The persistent pursuit's peak test of one thread's 8 columns w = 8t + q (q < 8, w < W): the
// window maximum pool_q over [w - nt, w + nt] (-inf outside the row) and the reference's test
// pool > Th^2, |pool - cf| < 1e-9 (max-pool equality), as bits q of the result.  The row cf is
// in LDS at pad(H + column) (one pad word per 32: the stride-8 reads below are conflict-free) and
// bm[b + kBmHalo] holds the maximum of block b (columns 8b .. 8b + 7).  For nt >= 7 every window of
// the 8 columns contains C = [8t + 7 - nt, 8t + nt]: its maximum comes from the whole blocks inside
// it and the two partial runs at its ends; column q adds head[q .. 6] (head[k] = column
// 8t - nt + k) and tail[0 .. q - 1] (tail[k] = column 8t + nt + 1 + k).  Smaller nt scans directly.
constexpr int kBmHalo = 17;   // blocks of -inf kept on each side of the row (nt <= 127)
__device__ __forceinline__ uint32_t peak_bits8(const float* cf, const float* bm, int t, int W, int nt, float th2) {
    const int H = kPeakHalo, c0 = 8 * t;
    float pool[8];
    if (nt >= 7) {
        const int bl = -((nt - 7) >> 3), bh = ((nt + 1) >> 3) - 1;
        const int ph = (nt - 7) & 7, pt = (nt + 1) & 7;
        float cm = -INFINITY;
        for (int b = t + bl; b <= t + bh; ++b) cm = fmaxf(cm, bm[b + kBmHalo]);
        const int s0 = c0 + 7 - nt, e1 = 8 * (t + bh + 1);
        for (int k = 0; k < ph; ++k) cm = fmaxf(cm, cf[pad32(H + s0 + k)]);
        for (int k = 0; k < pt; ++k) cm = fmaxf(cm, cf[pad32(H + e1 + k)]);
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

// LDS of the persistent pursuit for windows of W columns (floats/ints): the padded row with its
// halos and a -inf tail of 16, the block maxima with their halos, the covered-run prefix, imax,
// amplitude and the pass's list (column, template, amplitude)
__host__ __device__ constexpr int persist_cf_len(int W) { return pad_len(W + 2 * kPeakHalo + 16); }
__host__ __device__ constexpr int persist_bm_len(int W) { return (W + 7) / 8 + 2 * kBmHalo; }

template <int NV>
__global__ __launch_bounds__(kPT) void pursuit_persistent_kernel(PursuitArgs a, int first_pass, int max_passes) {
    // cf [persist_cf_len] | bm [persist_bm_len] | pfx [W + 1] | im [W] | am [W] | list col, tmp, amp [W]
    extern __shared__ float psm[];
    __shared__ int wsum[kPW];
    __shared__ int s_ok;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = (int)a.W, H = kPeakHalo;
    const int N = W + 2 * H;
    const int nbk = (W + 7) / 8;   // blocks of 8 columns, one per thread
    float* cf = psm;
    float* bm = cf + persist_cf_len(W);
    int* pfx = (int*)(bm + persist_bm_len(W));
    int* im = pfx + W + 1;
    float* am = (float*)(im + W);
    int* lcol = (int*)(am + W);
    int* ltmp = lcol + W;
    float* lamp = (float*)(ltmp + W);
    const bool lead = blockIdx.x == 0 && tid == 0;
    if (a.ts && tid == 0) a.ts[2 * blockIdx.x] = __builtin_amdgcn_s_memrealtime();   // measurement mode
    // this workgroup's columns: kColsB-aligned runs, G of them cover [0, W)
    const int Wc = (W + (int)gridDim.x * kColsB - 1) / ((int)gridDim.x * kColsB) * kColsB;
    const int a0 = min(W, (int)blockIdx.x * Wc), a1 = min(W, a0 + Wc);
    floatx4 nmv[NV], rnv[NV];
    load_norms<NV>(a, nmv, rnv);
    int par = first_pass & 1;
    PursuitArgs cur = a;
    use_row(cur, a, par);
    int base = 0, np = 0;
    bool done = false;
    if (first_pass == 0) {
        // first pass: every column scored, one wavefront per column over the whole launch
        for (int w = (int)blockIdx.x * kPW + wave; w < W; w += (int)gridDim.x * kPW) score_column<NV, true>(cur, w, lane, nmv, rnv);
    } else {
        // after first_pass passes of the per-pass kernels (their last pass committed as the peak
        // kernel commits it): own columns' Cfmax/imax as they left them, with their amplitudes
        base = a.counters[0] + a.counters[1];
        np = a.counters[4];
        done = a.counters[3] != 0;
        for (int w = a0 + tid; w < a1; w += kPT) {
            const int i = a.imax[w];
            store_score<true>(cur, w, a.cfmax[w], i, ld_dev(a.B + (int64_t)w * a.ldT + i), a.nm[i]);
        }
    }
    // constant -inf parts: the row's tail past N and the block halos
    if (tid < 16) cf[pad32(N + tid)] = -INFINITY;
    if (tid < kBmHalo) {
        bm[tid] = -INFINITY;
        bm[kBmHalo + nbk + tid] = -INFINITY;
    }
    int gen = 1;
    if (!grid_barrier(a.counters, gen++, &s_ok)) return;
    for (int pass = first_pass; pass < max_passes && !done; ++pass) {
        {
            // the row (with -inf halos) into LDS, every load in flight at once
            float v[kPersistLoads], b[kPersistLoads];
            int m[kPersistLoads];
#pragma unroll
            for (int r = 0; r < kPersistLoads; ++r) {
                const int w = tid + kPT * r - H;
                const bool in = w >= 0 && w < W;
                v[r] = in ? ld_dev(cur.cfmax + w) : -INFINITY;
                m[r] = in ? ld_dev(cur.imax + w) : 0;
                b[r] = in ? ld_dev(cur.amprow + w) : 0.f;
            }
#pragma unroll
            for (int r = 0; r < kPersistLoads; ++r) {
                const int i = tid + kPT * r, w = i - H;
                if (i < N) cf[pad32(i)] = v[r];
                if (w >= 0 && w < W) {
                    im[w] = m[r];
                    am[w] = b[r];
                }
            }
        }
        __syncthreads();
        if (tid < nbk) {
            float m = cf[pad32(H + 8 * tid)];
#pragma unroll
            for (int k = 1; k < 8; ++k) m = fmaxf(m, cf[pad32(H + 8 * tid + k)]);
            bm[kBmHalo + tid] = m;
        }
        __syncthreads();
        uint32_t bits = tid < nbk ? peak_bits8(cf, bm, tid, W, a.nt, a.th2) : 0u;
        const int cnt = __popc(bits);
        int inc = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int vv = __shfl_up(inc, off, 64);
            if (lane >= off) inc += vv;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int woff = 0, n = 0;
#pragma unroll
        for (int k = 0; k < kPW; ++k) {
            const int vv = wsum[k];
            if (k < wave) woff += vv;
            n += vv;
        }
        int pos = woff + inc - cnt;
        while (bits) {
            const int q = __ffs(bits) - 1;
            bits &= bits - 1;
            const int w = tid * 8 + q;
            const int t = im[w];
            const float amp = am[w];
            lcol[pos] = w;
            ltmp[pos] = t;
            lamp[pos] = amp;
            if (blockIdx.x == 0 && base + pos < a.cap) {
                st_dev(a.sp_col + base + pos, w);
                st_dev(a.sp_tmp + base + pos, t);
                st_dev(a.sp_amp + base + pos, amp);
            }
            ++pos;
        }
        if (n == 0) {
            if (lead) a.counters[3] = 1;
            break;
        }
        if (lead) {
            if (base + n > a.cap) a.counters[2] = 1;
            a.pass_off[np] = base;
            a.pass_off[np + 1] = min(base + n, a.cap);
        }
        ++np;
        __syncthreads();   // the pass's list is complete in LDS
        const int nu = max(0, min(n, a.cap - base));   // spikes stored, as update_kernel applies them
        PursuitArgs nxt = a;
        use_row(nxt, a, par ^ 1);
        const ListPlain lst{lcol, ltmp, lamp};
        const int K = covered_prefix(lst, nu, a.nt, W, pfx, wsum);
        carry_own_columns(nxt, lst, nu, a0, a1, cf, im, am);
        update_covered<NV>(nxt, lst, nu, pfx, K, nmv, rnv);
        base += n;
        par ^= 1;
        cur = nxt;
        if (pass + 1 < max_passes && !grid_barrier(a.counters, gen++, &s_ok)) return;
    }
    if (lead) {
        a.counters[0] = base;
        a.counters[1] = 0;
        a.counters[4] = np;
    }
    if (a.ts) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) a.ts[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
    }
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

template <int NVC>   // float4 channel chunks per lane: NVC*256 >= C
__global__ __launch_bounds__(256) void residual_kernel(PursuitArgs a) {
    __shared__ int s_first[kResMaxPasses], s_cnt[kResMaxPasses], s_off[kResMaxPasses + 1];
    __shared__ int l_col[kResCap], l_tmp[kResCap], l_pp[kResCap];   // l_pp = pass * 2 + parity
    __shared__ float l_amp[kResCap];
    __shared__ int l_apply[4][2 * kResMaxPasses];                    // per wavefront
    __shared__ int s_run[2];                                         // passes [s_run[0], s_run[1])
    if (a.counters[2]) return;                  // spike list overflowed: the call reports an error
    const int P = a.counters[4];                // passes that found spikes
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = a.nt / 2;
    const int64_t w0 = (int64_t)blockIdx.x * kResCols;
    const int lo = (int)(w0 - h), hi = (int)(w0 + kResCols - 1 + h);
    for (int p = tid; p < P; p += 256) {
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
        for (int p = p0 + tid; p < p1; p += 256) {
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
        for (int cw = wave; cw < kResCols; cw += 4) {
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

size_t persistent_lds(int64_t W) {
    const int w = (int)W;
    return sizeof(float) * (persist_cf_len(w) + persist_bm_len(w)) + sizeof(int) * (w + 1) +
           (sizeof(int) + sizeof(float)) * w + (2 * sizeof(int) + sizeof(float)) * w;
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

hipError_t persistent_max_grid(int T, int64_t W, int* grid) {
    const PersistentFn fn = persistent_fn(T);
    *grid = 0;
    if (!fn) return hipSuccess;   // more than 1024 templates: per-pass kernels only
    if (W < 1 || W > kPersistMaxW) return hipErrorInvalidValue;
    int dev = 0, cus = 0, per = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)persistent_lds(W));
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)fn, kPT, persistent_lds(W));
    if (e != hipSuccess) return e;
    *grid = per * cus;
    return hipSuccess;
}

hipError_t launch_pursuit_persistent(const PursuitArgs& a, int grid, int first_pass, int max_passes, hipStream_t s,
                                     int* used) {
    const PersistentFn fn = persistent_fn(a.T);
    if (!fn || a.W > kPersistMaxW || a.nt > kPeakHalo || grid < 1 || first_pass < 0 || first_pass >= max_passes ||
        !a.prow_cf || !a.prow_im || !a.prow_amp)
        return hipErrorInvalidValue;
    int cap = 0;
    hipError_t e = persistent_max_grid(a.T, a.W, &cap);
    if (e != hipSuccess) return e;
    if (cap < 1) return hipErrorInvalidConfiguration;
    if (used) *used = std::min(grid, cap);
    hipLaunchKernelGGL(fn, dim3(std::min(grid, cap)), dim3(kPT), persistent_lds(a.W), s, a, first_pass, max_passes);
    return hipGetLastError();
}

hipError_t launch_residual(const PursuitArgs& a, hipStream_t s) {
    if (a.C > 1024) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((a.W + kResCols - 1) / kResCols);
    switch ((a.C + 255) / 256) {
        case 1: hipLaunchKernelGGL(residual_kernel<1>, dim3(nb), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL(residual_kernel<2>, dim3(nb), dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL(residual_kernel<4>, dim3(nb), dim3(256), 0, s, a); break;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// PC features and positions (template_matching.py:50-51, postprocessing.py:33-46): one wavefront
// per spike.  The spike's nt x nC residual window is gathered into LDS with all loads in flight at
// once, then lane (j, k) forms sum_m X[w - nt/2 + m][iCC[j, iU[t]]] wPCA[k][m] + amp Ucc[j][t][k].
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void features_kernel(FeatureArgs a) {
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
    __syncthreads();
    for (int s0 = blockIdx.x * 4; s0 < n; s0 += gridDim.x * 4) {
        const int s = s0 + wave;
        if (s < n) {
            const int w = a.sp_col[s];
            const int t = a.sp_tmp[s];
            const float am = a.sp_amp[s];
            const int iu = a.iU[t];
            for (int e = lane; e < nt * nC; e += 64) {
                const int m = e / nC, j = e % nC;
                const int64_t ww = (int64_t)w - h + m;
                const int ch = a.iCC[j * a.C + iu];
                xwin[e] = (ww >= 0 && ww < a.W) ? a.X[ww * a.ldC + ch] : 0.f;
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
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
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
__global__ __launch_bounds__(256) void stage_spikes_kernel(StageArgs a) {
    int n = a.counters[0] + a.counters[1];
    n = min(n, a.cap);
    const int gid = blockIdx.x * 256 + threadIdx.x, stride = gridDim.x * 256;
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

hipError_t launch_stage_spikes(const StageArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(stage_spikes_kernel, dim3(64), dim3(256), 0, s, a);
    return hipGetLastError();
}

// counters[0..7] = 0 (a fill kernel, so that the sequence can be captured in a graph)
__global__ void reset_counters_kernel(int* c) {
    if (threadIdx.x < kCounters) c[threadIdx.x] = 0;
}

hipError_t launch_reset_counters(int* counters, hipStream_t s) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counters);
    return hipGetLastError();
}

}  // namespace lss
