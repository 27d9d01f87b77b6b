// Greedy matching pursuit over the score matrix B [W][T] (time-major: the T template scores of a
// sample are contiguous), Kilosort4 semantics (kilosort/template_matching.py:157-198):
//
//   Cf = relu(B)^2 / nm, zero in the first and last nt columns             (:159, :165-166)
//   Cfmax[w], imax[w] = max / first argmax over templates                  (:167)
//   peaks: maxpool_{2nt+1}(Cfmax)[w] > Th^2 and equal to Cfmax[w]          (:168-174)
//   amp = B[imax, w] / nm[imax]                                            (:188)
//   B[:, w-nt..w+nt] -= amp ctc[:, t, :];  Xres[:, w-nt/2..w+nt/2] -= amp wave_t
//   in two groups (even, then odd spike index of the pass)                 (:195-198)
//
// Scores are reduced with one wavefront per sample column (wavefront-shuffle argmax).  Peaks are
// compacted in ascending column order with a two-kernel block scan.  The update is a gather: the
// workgroup of spike s owns the columns of its window that no earlier spike's window covers and
// applies, per column, the last covering spike of each parity (even group first) - exactly the
// values the reference's gather-subtract-scatter leaves, including its last-write-wins result for
// windows of one group that overlap (tied peaks).  After updating a column it rescores it, so
// Cfmax/imax stay current without rescanning the whole matrix in later passes.
//
// FP contraction is off in this file: the products and differences are rounded separately, as
// the reference's separate tensor multiply and in-place subtraction are.
#pragma clang fp contract(off)
#include "lss_kernels.h"

namespace lss {

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

// one wavefront per column
__global__ __launch_bounds__(256) void score_all_kernel(const float* __restrict__ B,
                                                        const float* __restrict__ nm, int64_t W,
                                                        int T, int nt, float* __restrict__ cfmax,
                                                        int* __restrict__ imax) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= W) return;
    float best = -1.f;
    int bi = 0x7fffffff;
    if (w >= nt && w < W - nt) {
        const float* col = B + w * T;
        for (int t = lane; t < T; t += 64) argmax_merge(best, bi, score_of(col[t], nm[t]), t);
        wave_argmax(best, bi);
    } else {
        best = 0.f;
        bi = 0;
    }
    if (lane == 0) {
        cfmax[w] = best;
        imax[w] = bi;
    }
}

hipError_t launch_score_all(const PursuitArgs& a, hipStream_t s) {
    unsigned nb = (unsigned)((a.W + 3) / 4);
    hipLaunchKernelGGL(score_all_kernel, dim3(nb), dim3(256), 0, s, a.B, a.nm, a.W, a.T, a.nt,
                       a.cfmax, a.imax);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// peaks: 256 threads x 4 columns per block
// ------------------------------------------------------------------------------------------------
constexpr int kPeakHalo = 128;   // >= nt

__device__ __forceinline__ void peak_flags(const float* __restrict__ cfmax, int64_t W, int nt,
                                           float th2, float* lds, bool flags[4]) {
    const int64_t c0 = (int64_t)blockIdx.x * kPeakCols;
    for (int i = threadIdx.x; i < kPeakCols + 2 * kPeakHalo; i += 256) {
        int64_t w = c0 - kPeakHalo + i;
        lds[i] = (w >= 0 && w < W) ? cfmax[w] : -INFINITY;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int li = threadIdx.x * 4 + q;
        const int64_t w = c0 + li;
        bool f = false;
        if (w < W) {
            float cm = -INFINITY;
            for (int d = -nt; d <= nt; ++d) cm = fmaxf(cm, lds[kPeakHalo + li + d]);
            const float cf = lds[kPeakHalo + li];
            f = (cm > th2) && (fabsf(cm - cf) < 1e-9f);
        }
        flags[q] = f;
    }
}

__global__ __launch_bounds__(256) void peaks_count_kernel(PursuitArgs a) {
    __shared__ float lds[kPeakCols + 2 * kPeakHalo];
    __shared__ int wsum[4];
    // commit the previous pass (its update has completed: stream order)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.counters[0] += a.counters[1];
        a.counters[1] = 0;
    }
    bool f[4];
    peak_flags(a.cfmax, a.W, a.nt, a.th2, lds, f);
    int cnt = (int)f[0] + (int)f[1] + (int)f[2] + (int)f[3];
    for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) a.blockcnt[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void peaks_emit_kernel(PursuitArgs a) {
    __shared__ float lds[kPeakCols + 2 * kPeakHalo];
    __shared__ int wsum[4];
    __shared__ int s_off;
    if (threadIdx.x == 0) {
        int off = 0;
        for (int b = 0; b < (int)blockIdx.x; ++b) off += a.blockcnt[b];
        s_off = off;
        if (blockIdx.x == gridDim.x - 1) {
            int tot = off + a.blockcnt[blockIdx.x];
            a.counters[1] = tot;
            if (a.counters[0] + tot > a.cap) a.counters[2] = 1;
        }
    }
    bool f[4];
    peak_flags(a.cfmax, a.W, a.nt, a.th2, lds, f);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int cnt = (int)f[0] + (int)f[1] + (int)f[2] + (int)f[3];
    // inclusive wave scan
    int inc = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        int v = __shfl_up(inc, off, 64);
        if (lane >= off) inc += v;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int woff = 0;
    for (int q = 0; q < wave; ++q) woff += wsum[q];
    int pos = a.counters[0] + s_off + woff + inc - cnt;
    const int64_t c0 = (int64_t)blockIdx.x * kPeakCols;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (!f[q]) continue;
        if (pos < a.cap) {
            const int64_t w = c0 + threadIdx.x * 4 + q;
            const int t = a.imax[w];
            a.sp_col[pos] = (int)w;
            a.sp_tmp[pos] = t;
            a.sp_amp[pos] = a.B[w * a.T + t] / a.nm[t];
        }
        ++pos;
    }
}

hipError_t launch_peaks(const PursuitArgs& a, hipStream_t s) {
    unsigned nb = (unsigned)((a.W + kPeakCols - 1) / kPeakCols);
    hipLaunchKernelGGL(peaks_count_kernel, dim3(nb), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(peaks_emit_kernel, dim3(nb), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// update: one workgroup per spike of the pass (grid-stride), 4 wavefronts walk its columns
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void update_kernel(PursuitArgs a) {
    const int base = a.counters[0];
    int n = a.counters[1];
    if (base + n > a.cap) n = a.cap - base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nt = a.nt, h = nt / 2, T = a.T, C = a.C;
    const int L = 2 * nt + 1;
    const int* col = a.sp_col + base;
    const int* tmp = a.sp_tmp + base;
    const float* amp = a.sp_amp + base;
    for (int s = blockIdx.x; s < n; s += gridDim.x) {
        const int ws = col[s];
        // ---- score columns [ws-nt, ws+nt] not covered by spike s-1 ----
        int lo = ws - nt;
        if (s > 0) lo = max(lo, col[s - 1] + nt + 1);
        lo = max(lo, 0);
        const int hi = min((int64_t)ws + nt, a.W - 1);
        for (int w = lo + wave; w <= hi; w += 4) {
            int e = -1, o = -1;
            for (int j = s; j < n && col[j] - nt <= w; ++j) {
                if (j & 1) o = j; else e = j;
            }
            const bool rescore = (w >= nt) && (w < a.W - nt);
            float best = -1.f;
            int bi = 0x7fffffff;
            float* bcol = a.B + (int64_t)w * T;
            const float* ce = (e >= 0) ? a.ctc + ((int64_t)tmp[e] * L + (w - col[e] + nt)) * T : nullptr;
            const float* co = (o >= 0) ? a.ctc + ((int64_t)tmp[o] * L + (w - col[o] + nt)) * T : nullptr;
            const float ae = (e >= 0) ? amp[e] : 0.f;
            const float ao = (o >= 0) ? amp[o] : 0.f;
            for (int t = lane; t < T; t += 64) {
                float b = bcol[t];
                if (e >= 0) b = b - ae * ce[t];
                if (o >= 0) b = b - ao * co[t];
                bcol[t] = b;
                if (rescore) argmax_merge(best, bi, score_of(b, a.nm[t]), t);
            }
            if (rescore) {
                wave_argmax(best, bi);
                if (lane == 0) {
                    a.cfmax[w] = best;
                    a.imax[w] = bi;
                }
            }
        }
        // ---- residual columns [ws-h, ws+h] not covered by spike s-1 ----
        int xlo = ws - h;
        if (s > 0) xlo = max(xlo, col[s - 1] + h + 1);
        xlo = max(xlo, 0);
        const int xhi = min((int64_t)ws + h, a.W - 1);
        for (int w = xlo + wave; w <= xhi; w += 4) {
            int e = -1, o = -1;
            for (int j = s; j < n && col[j] - h <= w; ++j) {
                if (j & 1) o = j; else e = j;
            }
            float* xcol = a.X + (int64_t)w * C;
            const float* ve = (e >= 0) ? a.waves + ((int64_t)tmp[e] * nt + (w - col[e] + h)) * C : nullptr;
            const float* vo = (o >= 0) ? a.waves + ((int64_t)tmp[o] * nt + (w - col[o] + h)) * C : nullptr;
            const float ae = (e >= 0) ? amp[e] : 0.f;
            const float ao = (o >= 0) ? amp[o] : 0.f;
            for (int c = lane; c < C; c += 64) {
                float x = xcol[c];
                if (e >= 0) x = x - ae * ve[c];
                if (o >= 0) x = x - ao * vo[c];
                xcol[c] = x;
            }
        }
    }
}

hipError_t launch_update(const PursuitArgs& a, int n_new, hipStream_t s) {
    if (n_new <= 0) return hipSuccess;
    hipLaunchKernelGGL(update_kernel, dim3(n_new), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// PC features and positions: one wavefront per spike
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void features_kernel(FeatureArgs a) {
    __shared__ float sf[4][32 * 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = blockIdx.x * 4 + wave;
    const bool valid = s < a.n;
    const int h = a.nt / 2, K = a.K, nC = a.nC;
    int w = 0, t = 0, iu = 0;
    float am = 0.f;
    if (valid) {
        w = a.sp_col[s];
        t = a.sp_tmp[s];
        am = a.sp_amp[s];
        iu = a.iU[t];
        for (int p = lane; p < nC * K; p += 64) {
            const int j = p / K, k = p % K;
            const int ch = a.iCC[j * a.C + iu];
            float acc = 0.f;
            for (int m = 0; m < a.nt; ++m) {
                const int64_t ww = (int64_t)w - h + m;
                const float xv = (ww >= 0 && ww < a.W) ? a.X[ww * a.C + ch] : 0.f;
                acc = acc + xv * a.wpca[k * a.nt + m];
            }
            const float f = acc + am * a.Ucc[((int64_t)j * a.T + t) * K + k];
            sf[wave][p] = f;
            a.tF[(int64_t)s * nC * K + p] = f;
        }
    }
    __syncthreads();
    if (valid && lane == 0) {
        float tot = 0.f;
        for (int j = 0; j < nC; ++j) {
            float m = 0.f;
            for (int k = 0; k < K; ++k) m += sf[wave][j * K + k] * sf[wave][j * K + k];
            tot += m;
        }
        float xs = 0.f, ys = 0.f;
        for (int j = 0; j < nC; ++j) {
            float m = 0.f;
            for (int k = 0; k < K; ++k) m += sf[wave][j * K + k] * sf[wave][j * K + k];
            const int ch = a.iCC[j * a.C + iu];
            const float wgt = m / tot;
            xs += a.xc[ch] * wgt;
            ys += a.yc[ch] * wgt;
        }
        a.xs[s] = xs;
        a.ys[s] = ys;
    }
}

hipError_t launch_features(const FeatureArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (a.nC * a.K > 32 * 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(features_kernel, dim3((a.n + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace lss
