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

__device__ __forceinline__ float score_of(float b, float nmv) {
    float r = fmaxf(b, 0.f);
    return (r * r) / nmv;
}

// ------------------------------------------------------------------------------------------------
// full scan (first pass): one wavefront per column, float4 loads (ldT % 4 == 0)
// ------------------------------------------------------------------------------------------------
template <int NV>   // float4 vectors per lane: NV*256 >= T
__global__ __launch_bounds__(256) void score_all_kernel(PursuitArgs a) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= a.W) return;
    float best = 0.f;
    int bi = 0;
    if (w >= a.nt && w < a.W - a.nt) {
        best = -1.f;
        bi = 0x7fffffff;
        const float* col = a.B + w * a.ldT;
        floatx4 b[NV], n[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int t = 4 * lane + 256 * i;
            if (t < a.T) {
                b[i] = *(const floatx4*)(col + t);
                n[i] = *(const floatx4*)(a.nm + t);
            }
        }
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int t = 4 * lane + 256 * i;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (t + e < a.T) argmax_merge(best, bi, score_of(b[i][e], n[i][e]), t + e);
        }
        wave_argmax(best, bi);
    }
    if (lane == 0) {
        a.cfmax[w] = best;
        a.imax[w] = bi;
    }
}

// ------------------------------------------------------------------------------------------------
// peaks.  A column is a peak iff pool > Th^2 and |pool - cf| < 1e-9 (pool = window max).  If any
// window value v has fl(v - cf) >= 1e-9 then fl(pool - cf) >= fl(v - cf) (rounding is monotone),
// so the column is not a peak: the scan stops at the first such v, nearest neighbours first.  A
// column with cf <= Th^2 - 2e-9 cannot reach pool > Th^2 with pool - cf < 1e-9 and is skipped.
// Columns that survive get the literal test on the full window maximum.
//
// peaks_flags_kernel: one thread per column over the whole chip -> flag bytes.
// peaks_emit_kernel: one workgroup compacts the flags in ascending column order.
// ------------------------------------------------------------------------------------------------
constexpr int kFlagCols = 256;
constexpr int kPeakHalo = 128;                       // >= nt
constexpr int kEmitThreads = 1024;
constexpr int kEmitPer = 64;                         // columns per thread per emit tile
constexpr int kEmitTile = kEmitThreads * kEmitPer;   // 65536

__global__ __launch_bounds__(kFlagCols) void peaks_flags_kernel(PursuitArgs a) {
    __shared__ float lds[kFlagCols + 2 * kPeakHalo];
    if (a.counters[3]) return;   // converged earlier: nothing can change
    const int tid = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * kFlagCols;
    float v[(kFlagCols + 2 * kPeakHalo) / kFlagCols];
#pragma unroll
    for (int r = 0; r < (kFlagCols + 2 * kPeakHalo) / kFlagCols; ++r) {
        const int64_t w = c0 - kPeakHalo + tid + r * kFlagCols;
        v[r] = (w >= 0 && w < a.W) ? a.cfmax[w] : -INFINITY;
    }
#pragma unroll
    for (int r = 0; r < (kFlagCols + 2 * kPeakHalo) / kFlagCols; ++r) lds[tid + r * kFlagCols] = v[r];
    __syncthreads();
    const int64_t w = c0 + tid;
    if (w >= a.W) return;
    const float* p = lds + kPeakHalo + tid;
    const float cf = p[0];
    const float th2 = a.th2;
    bool f = false;
    if ((double)cf > (double)th2 - 2e-9 && !(p[-1] - cf >= 1e-9f) && !(p[1] - cf >= 1e-9f)) {
        float cm = fmaxf(cf, fmaxf(p[-1], p[1]));
        bool dominated = false;
        for (int d = 2; d <= a.nt; ++d) {
            const float vl = p[-d], vr = p[d];
            if (vl - cf >= 1e-9f || vr - cf >= 1e-9f) { dominated = true; break; }
            cm = fmaxf(cm, fmaxf(vl, vr));
        }
        f = !dominated && cm > th2 && fabsf(cm - cf) < 1e-9f;
    }
    a.flags[w] = f ? 1 : 0;
}

__global__ __launch_bounds__(kEmitThreads) void peaks_emit_kernel(PursuitArgs a) {
    __shared__ int wsum[kEmitThreads / 64];
    __shared__ int s_base, s_run;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        // commit the previous pass (its update has completed: stream order)
        const int base = a.counters[0] + a.counters[1];
        a.counters[0] = base;
        a.counters[1] = 0;
        s_base = base;
        s_run = 0;
    }
    __syncthreads();
    if (a.counters[3]) return;
    for (int64_t c0 = 0; c0 < a.W; c0 += kEmitTile) {
        // kEmitPer flag bytes per thread as four 16-byte loads (flags has kEmitTile padding)
        const uint4* fp = (const uint4*)(a.flags + c0 + (int64_t)tid * kEmitPer);
        uint4 q[kEmitPer / 16];
#pragma unroll
        for (int j = 0; j < kEmitPer / 16; ++j) q[j] = fp[j];
        uint64_t bits = 0;
#pragma unroll
        for (int j = 0; j < kEmitPer / 16; ++j) {
            const uint32_t wv[4] = {q[j].x, q[j].y, q[j].z, q[j].w};
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if ((wv[k] >> (8 * b)) & 0xffu) bits |= 1ull << (16 * j + 4 * k + b);
        }
        // bytes past W may hold flags of an earlier, longer window
        const int64_t valid = a.W - (c0 + (int64_t)tid * kEmitPer);
        if (valid <= 0) bits = 0;
        else if (valid < 64) bits &= (1ull << valid) - 1;
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
        for (int k = 0; k < kEmitThreads / 64; ++k) {
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

// ------------------------------------------------------------------------------------------------
// update: work items (spike s, chunk of kColsB of its window's score columns)
// ------------------------------------------------------------------------------------------------
constexpr int kColsB = 8;     // score columns per work item (2 per wavefront)
constexpr int kUpdateGrid = 8192;   // workgroups per update launch (grid-stride over work items)

// Spikes of the pass covering column w within +-reach, starting at spike s (whose window covers
// w): the last even-indexed (e) and odd-indexed (o) ones.  Spikes are in ascending column order,
// so the covering ones are s, s+1, ...; c1, c2 are col[s+1], col[s+2] (INT_MAX past the end).
__device__ __forceinline__ void covering(const int* col, int s, int n, int c1, int c2, int w, int reach,
                                         int& e, int& o) {
    int last = s;
    if (c1 - reach <= w) {
        last = s + 1;
        if (c2 - reach <= w) {
            last = s + 2;
            while (last + 1 < n && col[last + 1] - reach <= w) ++last;   // tied peaks only
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

template <int NV>   // float4 vectors per lane over templates: NV*256 >= T
__device__ __forceinline__ void update_score_columns(const PursuitArgs& a, const int* col, const int* tmp,
                                                     const float* amp, int s, int n, int c1, int c2, int c0,
                                                     int lo, int w1, const floatx4* nmv) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nt = a.nt, L = 2 * nt + 1;
    // two columns per wavefront, loads of both issued before any use
    const int wa = c0 + 2 * wave, wb = wa + 1;
    const bool va = wa >= lo && wa <= w1, vb = wb >= lo && wb <= w1;
    int ea, oa, eb, ob;
    covering(col, s, n, c1, c2, wa, nt, ea, oa);
    covering(col, s, n, c1, c2, wb, nt, eb, ob);
    floatx4 ba[NV], bb[NV], cea[NV], coa[NV], ceb[NV], cob[NV];
    float* pa = a.B + (int64_t)wa * a.ldT;
    float* pb = a.B + (int64_t)wb * a.ldT;
    const float* qea = (ea >= 0) ? a.ctc + ((int64_t)tmp[ea] * L + (wa - col[ea] + nt)) * a.ldT : nullptr;
    const float* qoa = (oa >= 0) ? a.ctc + ((int64_t)tmp[oa] * L + (wa - col[oa] + nt)) * a.ldT : nullptr;
    const float* qeb = (eb >= 0) ? a.ctc + ((int64_t)tmp[eb] * L + (wb - col[eb] + nt)) * a.ldT : nullptr;
    const float* qob = (ob >= 0) ? a.ctc + ((int64_t)tmp[ob] * L + (wb - col[ob] + nt)) * a.ldT : nullptr;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int t = 4 * lane + 256 * i;
        if (t < a.T) {
            if (va) {
                ba[i] = *(const floatx4*)(pa + t);
                if (qea) cea[i] = *(const floatx4*)(qea + t);
                if (qoa) coa[i] = *(const floatx4*)(qoa + t);
            }
            if (vb) {
                bb[i] = *(const floatx4*)(pb + t);
                if (qeb) ceb[i] = *(const floatx4*)(qeb + t);
                if (qob) cob[i] = *(const floatx4*)(qob + t);
            }
        }
    }
    const float aea = (ea >= 0) ? amp[ea] : 0.f, aoa = (oa >= 0) ? amp[oa] : 0.f;
    const float aeb = (eb >= 0) ? amp[eb] : 0.f, aob = (ob >= 0) ? amp[ob] : 0.f;
    const bool ra = va && wa >= nt && wa < a.W - nt;
    const bool rb = vb && wb >= nt && wb < a.W - nt;
    float besta = -1.f, bestb = -1.f;
    int ia = 0x7fffffff, ib = 0x7fffffff;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int t = 4 * lane + 256 * i;
        if (t < a.T) {
            if (va) {
                floatx4 x = ba[i];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float v = x[e];
                    if (qea) v = v - aea * cea[i][e];
                    if (qoa) v = v - aoa * coa[i][e];
                    x[e] = v;
                    if (ra && t + e < a.T) argmax_merge(besta, ia, score_of(v, nmv[i][e]), t + e);
                }
                *(floatx4*)(pa + t) = x;
            }
            if (vb) {
                floatx4 x = bb[i];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float v = x[e];
                    if (qeb) v = v - aeb * ceb[i][e];
                    if (qob) v = v - aob * cob[i][e];
                    x[e] = v;
                    if (rb && t + e < a.T) argmax_merge(bestb, ib, score_of(v, nmv[i][e]), t + e);
                }
                *(floatx4*)(pb + t) = x;
            }
        }
    }
    if (ra) {
        wave_argmax(besta, ia);
        if (lane == 0) { a.cfmax[wa] = besta; a.imax[wa] = ia; }
    }
    if (rb) {
        wave_argmax(bestb, ib);
        if (lane == 0) { a.cfmax[wb] = bestb; a.imax[wb] = ib; }
    }
}

template <int NV>
__global__ __launch_bounds__(256) void update_kernel(PursuitArgs a) {
    const int base = a.counters[0];
    int n = a.counters[1];
    if (n == 0) return;
    if (base + n > a.cap) n = a.cap - base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nt = a.nt, h = nt / 2;
    const int* col = a.sp_col + base;
    const int* tmp = a.sp_tmp + base;
    const float* amp = a.sp_amp + base;
    const int nB = (2 * nt + 1 + kColsB - 1) / kColsB;
    const int per = nB;
    floatx4 nmv[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int t = 4 * lane + 256 * i;
        if (t < a.T) nmv[i] = *(const floatx4*)(a.nm + t);
    }
    for (int64_t q = blockIdx.x; q < (int64_t)n * per; q += gridDim.x) {
        const int s = (int)(q / per), sub = (int)(q % per);
        // the spike's column and its neighbours' (independent loads, issued together)
        const int ws = col[s];
        const int cp = (s > 0) ? col[s - 1] : INT_MIN / 2;
        const int c1 = (s + 1 < n) ? col[s + 1] : INT_MAX / 2;
        const int c2 = (s + 2 < n) ? col[s + 2] : INT_MAX / 2;
        if (sub < nB) {
            int lo = ws - nt;
            if (s > 0) lo = max(lo, cp + nt + 1);
            lo = max(lo, 0);
            const int hi = (int)min((int64_t)ws + nt, a.W - 1);
            const int c0 = ws - nt + sub * kColsB;
            const int w1 = min(c0 + kColsB - 1, hi);
            if (max(c0, lo) <= w1) update_score_columns<NV>(a, col, tmp, amp, s, n, c1, c2, c0, lo, w1, nmv);
        }
    }
}

// ------------------------------------------------------------------------------------------------
// residual, once after the last pass.  A workgroup owns 64 columns; it finds, per pass, the spikes
// whose +-nt/2 window reaches its columns (binary search in the pass's column-sorted list), lists
// them in pursuit order in LDS, and each wavefront walks that list for its columns applying, pass
// by pass, the last even then the last odd covering spike:
//   X[w][c] -= amp * waves[t][w - col + nt/2][c]
// which is the residual the reference's per-pass updates leave (template_matching.py:197).
// ------------------------------------------------------------------------------------------------
constexpr int kResCols = 16;
constexpr int kResCap = 1024;       // listed spikes per workgroup
constexpr int kResMaxPasses = 256;
constexpr int kResBatch = 8;        // wave-row loads in flight per column

template <int NVC>   // float4 channel chunks per lane: NVC*256 >= C
__global__ __launch_bounds__(256) void residual_kernel(PursuitArgs a) {
    __shared__ int s_first[kResMaxPasses], s_cnt[kResMaxPasses], s_off[kResMaxPasses + 1];
    __shared__ int l_col[kResCap], l_tmp[kResCap], l_pp[kResCap];   // l_pp = pass * 2 + parity
    __shared__ float l_amp[kResCap];
    __shared__ int l_apply[4][2 * kResMaxPasses];                    // per wavefront
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
    if (tid == 0) {
        int acc = 0;
        for (int p = 0; p < P; ++p) { s_off[p] = acc; acc += s_cnt[p]; }
        s_off[P] = acc;
        if (acc > kResCap) a.counters[5] = 1;
    }
    __syncthreads();
    const int total = min(s_off[P], kResCap);
    for (int p = tid; p < P; p += 256) {
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
    if (s_off[P] > kResCap) return;
    constexpr int nvc = NVC;
    int* apply = l_apply[wave];
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
}

// ------------------------------------------------------------------------------------------------
hipError_t launch_score_all(const PursuitArgs& a, hipStream_t s) {
    unsigned nb = (unsigned)((a.W + 3) / 4);
    const int nv = (a.T + 255) / 256;
    switch (nv) {
#define LSS_SA(V) case V: hipLaunchKernelGGL(score_all_kernel<V>, dim3(nb), dim3(256), 0, s, a); break;
        LSS_SA(1) LSS_SA(2) LSS_SA(3) LSS_SA(4)
#undef LSS_SA
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_peaks(const PursuitArgs& a, hipStream_t s) {
    if (a.nt > kPeakHalo) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((a.W + kFlagCols - 1) / kFlagCols);
    hipLaunchKernelGGL(peaks_flags_kernel, dim3(nb), dim3(kFlagCols), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(peaks_emit_kernel, dim3(1), dim3(kEmitThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_update(const PursuitArgs& a, int grid, hipStream_t s) {
    const int nv = (a.T + 255) / 256;
    if (grid < 1) grid = kUpdateGrid;
    switch (nv) {
#define LSS_UP(V) case V: hipLaunchKernelGGL(update_kernel<V>, dim3(grid), dim3(256), 0, s, a); break;
        LSS_UP(1) LSS_UP(2) LSS_UP(3) LSS_UP(4)
#undef LSS_UP
        default: return hipErrorInvalidValue;
    }
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
    extern __shared__ float fsm[];   // wpca [K*nt] | per wavefront: window [nt][nC], features [nC*K]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int n = a.counters ? a.counters[0] + a.counters[1] : a.n;
    if (n > a.cap) n = a.cap;
    const int h = a.nt / 2, K = a.K, nC = a.nC, nt = a.nt;
    float* wp = fsm;
    float* xwin = fsm + K * nt + wave * (nt * nC + nC * K);
    float* sf = xwin + nt * nC;
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
            if (lane == 0) {
                float tot = 0.f;
                for (int j = 0; j < nC; ++j) {
                    float m = 0.f;
                    for (int k = 0; k < K; ++k) m += sf[j * K + k] * sf[j * K + k];
                    tot += m;
                }
                float xs = 0.f, ys = 0.f;
                for (int j = 0; j < nC; ++j) {
                    float m = 0.f;
                    for (int k = 0; k < K; ++k) m += sf[j * K + k] * sf[j * K + k];
                    const int ch = a.iCC[j * a.C + iu];
                    const float wgt = m / tot;
                    xs += a.xc[ch] * wgt;
                    ys += a.yc[ch] * wgt;
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
    const size_t lds = sizeof(float) * ((size_t)a.K * a.nt + 4 * ((size_t)a.nt * a.nC + a.nC * a.K));
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)features_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(features_kernel, dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

// counters[0..7] = 0 (a fill kernel, so that the sequence can be captured in a graph)
__global__ void reset_counters_kernel(int* c) {
    if (threadIdx.x < 8) c[threadIdx.x] = 0;
}

hipError_t launch_reset_counters(int* counters, hipStream_t s) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counters);
    return hipGetLastError();
}

}  // namespace lss
