// Template refinement: the residual energy of a fixed spike set and its adjoint.
//
//   R[w][c] = X[w][c] - sum_s amp_s * wave_{t_s}[w - col_s + nt/2][c]      (|w - col_s| <= nt/2,
//             wave_t[m][c] = sum_k U[t][k][c] wPCA[k][m], the update of template_matching.py:197;
//             contributions add, columns outside the window are dropped)
//   L = sum R^2
//   G_t[m][c] = sum_{s: t_s = t} amp_s R[col_s - nt/2 + m][c]   (spike-aligned residual windows)
//   dL/dU[t][k][c] = -2 sum_m wPCA[k][m] G_t[m][c]
//   dL/dwPCA[k][m] = -2 sum_t sum_c U[t][k][c] G_t[m][c]
//
// The residual is recomputed here from X (nothing of the forward pass is stored).  Every sum runs
// in a fixed order (spikes in list order, templates in index order), so results are
// deterministic.  The spike list must be sorted by column (ascending; ties in any order).
#include "lss_kernels.h"

namespace lss {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kRefCols = 16;     // residual columns per workgroup
constexpr int kRefList = 2048;   // spikes of one template held in LDS at a time
constexpr int kRefRows = 4;      // template rows accumulated together by one wavefront
constexpr int kRefMaxK = 8;

// first s in [0, n) with col[s] >= v
__device__ __forceinline__ int lower_bound_col(const int* col, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------------------------------------------
// residual: a workgroup owns kRefCols columns; each wavefront one column at a time, lanes over
// channels in groups of four.  Block partial sums of R^2 in FP64.
// ------------------------------------------------------------------------------------------------
template <int NVC>
__global__ __launch_bounds__(256) void refine_residual_kernel(RefineArgs a) {
    __shared__ int s_rng[2];
    __shared__ double s_red[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = a.nt / 2;
    const int64_t w0 = (int64_t)blockIdx.x * kRefCols;
    if (threadIdx.x == 0) {
        s_rng[0] = lower_bound_col(a.col, a.n, (int)w0 - h);
        s_rng[1] = lower_bound_col(a.col, a.n, (int)w0 + kRefCols + h);
    }
    __syncthreads();
    const int lo = s_rng[0], hi = s_rng[1];
    const bool vec_x = a.vec_x != 0;
    double e = 0.0;
    for (int j = wave; j < kRefCols; j += 4) {
        const int64_t w = w0 + j;
        if (w >= a.W) break;
        floatx4 x[NVC];
#pragma unroll
        for (int v = 0; v < NVC; ++v) {
            const int c = 4 * lane + 256 * v;
            const float* xp = a.X + w * a.ldx + c;
            if (c + 3 < a.C && vec_x) {
                x[v] = *(const floatx4*)xp;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) x[v][r] = (c + r < a.C) ? xp[r] : 0.f;
            }
        }
        for (int s = lo; s < hi; ++s) {
            const int d = (int)(w - a.col[s]) + h;
            if (d < 0 || d >= a.nt) continue;
            const int t = a.tmp[s];
            if (t < 0 || t >= a.T) {
                if (lane == 0) a.flag[0] = 1;
                continue;
            }
            const float am = a.amp[s];
            const float* wr = a.waves + ((int64_t)t * a.nt + d) * a.ldC;
#pragma unroll
            for (int v = 0; v < NVC; ++v) {
                const int c = 4 * lane + 256 * v;
                if (c < a.C) x[v] -= am * *(const floatx4*)(wr + c);
            }
        }
#pragma unroll
        for (int v = 0; v < NVC; ++v) {
            const int c = 4 * lane + 256 * v;
            if (c < a.C) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (c + r >= a.C) x[v][r] = 0.f;
                *(floatx4*)(a.R + w * a.ldC + c) = x[v];
#pragma unroll
                for (int r = 0; r < 4; ++r) e += (double)x[v][r] * (double)x[v][r];
            }
        }
    }
    e = wave_sum_d(e);
    if (lane == 0) s_red[wave] = e;
    __syncthreads();
    if (threadIdx.x == 0) a.part[blockIdx.x] = ((s_red[0] + s_red[1]) + s_red[2]) + s_red[3];
}

// loss = sum of the block partials, in block order (one workgroup)
__global__ __launch_bounds__(256) void refine_loss_kernel(const double* part, int nb, double* loss) {
    __shared__ double s_red[256];
    double e = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) e += part[i];
    s_red[threadIdx.x] = e;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s_red[threadIdx.x] += s_red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = s_red[0];
}

// ------------------------------------------------------------------------------------------------
// adjoint: one workgroup per (template t, chunk of kRefChunk rows m).  The template's spikes are
// collected from the list in order (ordered compaction, 1024 entries per step) into LDS,
// kRefList at a time; wavefront v accumulates rows m0 + 4v .. m0 + 4v + 3 of
//   G_t[m][c] = sum_s amp_s R[col_s - nt/2 + m][c]
// over the spikes (lanes over channels; SU spikes with all their loads in flight per step, summed
// in list order).  The finished rows go to LDS, from where the workgroup forms its share of dU
// (partial over its rows, summed over chunks by refine_du_kernel) and the rows of
// dW[t][k][m] = sum_c U[t][k][c] G_t[m][c].
// ------------------------------------------------------------------------------------------------
constexpr int kRefChunk = 16;

template <int NVC, int SU>
__global__ __launch_bounds__(256) void refine_adjoint_kernel(RefineArgs a) {
    extern __shared__ float rsm[];   // G rows [kRefChunk][ldC] | U_t [K][ldC] | list col, amp [kRefList]
    __shared__ int s_wcnt[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int K = a.K, nt = a.nt, h = nt / 2;
    const int nch = (nt + kRefChunk - 1) / kRefChunk;
    const int t = blockIdx.x / nch, ch = blockIdx.x % nch;
    const int m0 = ch * kRefChunk;
    const int ldc = (int)a.ldC;
    float* g = rsm;
    float* su = g + kRefChunk * ldc;
    int* lcol = (int*)(su + K * ldc);
    float* lamp = (float*)(lcol + kRefList);
    for (int e = tid; e < K * ldc; e += 256) {
        const int k = e / ldc, c = e % ldc;
        su[e] = (c < a.C) ? a.U[(int64_t)t * K * a.C + (int64_t)k * a.C + c] : 0.f;
    }
    floatx4 acc[4][NVC];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
#pragma unroll
        for (int v = 0; v < NVC; ++v) acc[rr][v] = floatx4{0.f, 0.f, 0.f, 0.f};
    int cnt = 0;
    for (int i0 = 0; i0 < a.n; i0 += 1024) {
        // four consecutive list entries per thread: ordered compaction of this template's spikes
        int hits = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = i0 + 4 * tid + e;
            if (i < a.n && a.tmp[i] == t) hits |= 1 << e;
        }
        const int my = __popc(hits);
        int inc = my;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int vv = __shfl_up(inc, off, 64);
            if (lane >= off) inc += vv;
        }
        if (lane == 63) s_wcnt[wave] = inc;
        __syncthreads();
        int off = cnt, tot = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < wave) off += s_wcnt[k];
            tot += s_wcnt[k];
        }
        int p = off + inc - my;
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (hits & (1 << e)) {
                const int i = i0 + 4 * tid + e;
                lcol[p] = a.col[i];
                lamp[p] = a.amp[i];
                ++p;
            }
        cnt += tot;
        __syncthreads();
        if (cnt <= kRefList - 1024 && i0 + 1024 < a.n) continue;
        // accumulate this segment, SU spikes per step
        for (int j = 0; j < cnt; j += SU) {
            floatx4 r[SU][4][NVC];
            float am[SU];
#pragma unroll
            for (int q = 0; q < SU; ++q) {
                const bool valid = j + q < cnt;
                const int cs = valid ? lcol[j + q] : 0;
                am[q] = valid ? lamp[j + q] : 0.f;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int m = m0 + 4 * wave + rr;
                    const int64_t w = (int64_t)cs - h + m;
                    const bool ok = valid && m < nt && w >= 0 && w < a.W;
                    const float* rp = a.R + (ok ? w : 0) * a.ldC;
#pragma unroll
                    for (int v = 0; v < NVC; ++v) {
                        const int c = 4 * lane + 256 * v;
                        r[q][rr][v] = *(const floatx4*)(rp + min(c, ldc - 4));
                        if (!ok || c >= a.C) r[q][rr][v] = floatx4{0.f, 0.f, 0.f, 0.f};
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < SU; ++q)
#pragma unroll
                for (int rr = 0; rr < 4; ++rr)
#pragma unroll
                    for (int v = 0; v < NVC; ++v) acc[rr][v] += am[q] * r[q][rr][v];
        }
        cnt = 0;
        __syncthreads();
    }
    // G rows to LDS
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
#pragma unroll
        for (int v = 0; v < NVC; ++v) {
            const int c = 4 * lane + 256 * v;
            if (c < ldc) *(floatx4*)(g + (4 * wave + rr) * ldc + c) = acc[rr][v];
        }
    __syncthreads();
    const int mrows = min(kRefChunk, nt - m0);
    // dU share: sum over this chunk's rows of wPCA[k][m] G[m][c]
    float* dup = a.dupart + ((int64_t)t * nch + ch) * K * a.C;
    for (int e = tid; e < K * a.C; e += 256) {
        const int k = e / a.C, c = e % a.C;
        float sacc = 0.f;
        for (int m = 0; m < mrows; ++m) sacc += a.wpca[k * nt + m0 + m] * g[m * ldc + c];
        dup[e] = sacc;
    }
    // dW rows: one wavefront per (m, k) pair, lanes over channels
    for (int pr = wave; pr < mrows * K; pr += 4) {
        const int m = pr / K, k = pr % K;
        float sacc = 0.f;
        for (int c = lane; c < a.C; c += 64) sacc += su[k * ldc + c] * g[m * ldc + c];
        sacc = wave_sum_f(sacc);
        if (lane == 0) a.dwpart[((int64_t)t * K + k) * nt + m0 + m] = sacc;
    }
}

// gU[t][k][c] = -2 sum over row chunks of the dU shares (chunk order)
__global__ __launch_bounds__(256) void refine_du_kernel(const float* dupart, int T, int nch, int KC, float* gU) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)T * KC) return;
    const int64_t t = e / KC, r = e % KC;
    float sacc = 0.f;
    for (int ch = 0; ch < nch; ++ch) sacc += dupart[(t * nch + ch) * KC + r];
    gU[e] = -2.f * sacc;
}

// gW[k][m] = -2 sum_t dwpart[t][k][m]: one workgroup per entry, strided partial sums over t in
// index order, then a fixed tree (FP64)
__global__ __launch_bounds__(256) void refine_dw_kernel(const float* dwpart, int T, int KM, float* gW) {
    __shared__ double s_red[256];
    const int e = blockIdx.x;
    double sacc = 0.0;
    for (int t = threadIdx.x; t < T; t += 256) sacc += (double)dwpart[(int64_t)t * KM + e];
    s_red[threadIdx.x] = sacc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s_red[threadIdx.x] += s_red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) gW[e] = (float)(-2.0 * s_red[0]);
}

// out[i] = (float)(x[i] - alpha * y[i]) (alpha in FP64)
__global__ __launch_bounds__(256) void step_kernel(float* out, const float* x, const float* y, double alpha, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        out[i] = (float)((double)x[i] - alpha * (double)y[i]);
}

// part[b] = sum of squares of x over block b's grid-stride share (FP64), then one workgroup sums
// the partials in order
constexpr int kSumBlocks = 256;
__global__ __launch_bounds__(256) void sumsq_kernel(const float* x, int64_t n, double* part) {
    __shared__ double s_red[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kSumBlocks * 256)
        acc += (double)x[i] * (double)x[i];
    s_red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s_red[threadIdx.x] += s_red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s_red[0];
}

hipError_t launch_step(float* out, const float* x, const float* y, double alpha, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(step_kernel, dim3(1024), dim3(256), 0, s, out, x, y, alpha, n);
    return hipGetLastError();
}

// out[0] += sum of squares of x (out must be initialised); part: kSumBlocks doubles of scratch
hipError_t launch_sumsq(const float* x, int64_t n, double* part, double* out, hipStream_t s) {
    hipLaunchKernelGGL(sumsq_kernel, dim3(kSumBlocks), dim3(256), 0, s, x, n, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(refine_loss_kernel, dim3(1), dim3(256), 0, s, part, kSumBlocks, out);
    return hipGetLastError();
}

__global__ void fill_i32_kernel(int* p, int n, int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_i32(int* p, int n, int v, hipStream_t s) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

size_t refine_adjoint_lds(const RefineArgs& a) {
    return sizeof(float) * (kRefChunk + a.K) * a.ldC + (sizeof(int) + sizeof(float)) * kRefList;
}

int refine_chunks(int nt) { return (nt + kRefChunk - 1) / kRefChunk; }

hipError_t launch_refine_residual(const RefineArgs& a, hipStream_t s) {
    if (a.C > 1024 || a.nt > 127 || a.W < 1) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((a.W + kRefCols - 1) / kRefCols);
    switch ((a.C + 255) / 256) {
        case 1: hipLaunchKernelGGL(refine_residual_kernel<1>, dim3(nb), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL(refine_residual_kernel<2>, dim3(nb), dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL(refine_residual_kernel<4>, dim3(nb), dim3(256), 0, s, a); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(refine_loss_kernel, dim3(1), dim3(256), 0, s, a.part, (int)nb, a.loss);
    return hipGetLastError();
}

int refine_blocks(int64_t W) { return (int)((W + kRefCols - 1) / kRefCols); }

template <int NVC, int SU>
hipError_t adjoint_launch(const RefineArgs& a, size_t lds, unsigned grid, hipStream_t s) {
    hipError_t e = hipFuncSetAttribute((const void*)refine_adjoint_kernel<NVC, SU>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((refine_adjoint_kernel<NVC, SU>), dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_refine_adjoint(const RefineArgs& a, hipStream_t s) {
    if (a.C > 1024 || a.K > kRefMaxK || a.nt > 127 || a.T < 1) return hipErrorInvalidValue;
    const size_t lds = refine_adjoint_lds(a);
    const int nch = refine_chunks(a.nt);
    const unsigned grid = (unsigned)(a.T * nch);
    hipError_t e;
    switch ((a.C + 255) / 256) {
        case 1: e = adjoint_launch<1, 4>(a, lds, grid, s); break;
        case 2: e = adjoint_launch<2, 4>(a, lds, grid, s); break;
        default: e = adjoint_launch<4, 2>(a, lds, grid, s); break;
    }
    if (e != hipSuccess) return e;
    const int KC = a.K * a.C;
    hipLaunchKernelGGL(refine_du_kernel, dim3((unsigned)(((int64_t)a.T * KC + 255) / 256)), dim3(256), 0, s,
                       a.dupart, a.T, nch, KC, a.gU);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int KM = a.K * a.nt;
    hipLaunchKernelGGL(refine_dw_kernel, dim3(KM), dim3(256), 0, s, a.dwpart, a.T, KM, a.gW);
    return hipGetLastError();
}

}  // namespace lss
