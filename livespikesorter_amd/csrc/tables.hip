// Load-time tables derived from the templates, accumulated in FP64 and stored as FP32:
//   nm[t]          = sum_{k,c} U[t,k,c]^2                            (template_matching.py:141)
//   waves[t][m][c] = sum_k U[t,k,c] wPCA[k,m]                        (the residual update's wave,
//                                                                     template_matching.py:197)
//   ctc[j][l][i]   = sum_{k,m} UtU[i,j,k,m] WtW[k,m,l]               (template_matching.py:110-121)
// with UtU[i,j,k,m] = sum_c U[i,k,c] U[j,m,c] and WtW[k,m,l] = sum_tau W[k][nt-l+tau] W[m][tau]
// (conv1d(wPCA, wPCA, padding=nt) flipped).  The ctc layout puts the T scores of one (template,
// lag) pair contiguously, which is the row the pursuit update streams.
#include "lss_kernels.h"

namespace lss {

__global__ __launch_bounds__(256) void nm_kernel(const float* __restrict__ U, int64_t ldu, int KC,
                                                 float* __restrict__ nm) {
    __shared__ double red[256];
    const float* u = U + (int64_t)blockIdx.x * ldu;
    double acc = 0.0;
    for (int i = threadIdx.x; i < KC; i += 256) acc += (double)u[i] * (double)u[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) nm[blockIdx.x] = (float)red[0];
}

__global__ __launch_bounds__(256) void waves_kernel(const float* __restrict__ U, int64_t ldu, const float* __restrict__ wpca,
                                                    int T, int K, int C, int nt, float* __restrict__ waves,
                                                    int64_t ldC) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)T * nt * C;
    if (idx >= total) return;
    const int c = (int)(idx % C);
    const int m = (int)((idx / C) % nt);
    const int t = (int)(idx / ((int64_t)C * nt));
    double acc = 0.0;
    for (int k = 0; k < K; ++k) acc += (double)U[(int64_t)t * ldu + (int64_t)k * C + c] * (double)wpca[k * nt + m];
    waves[((int64_t)t * nt + m) * ldC + c] = (float)acc;
}

// G[(i,k)][(j,m)] = sum_c U[i,k,c] U[j,m,c]: 16x16 FP64 tiles staged in LDS; row r = i*K+k of U
// starts at (r/K)*ldu + (r%K)*C.
__global__ __launch_bounds__(256) void gram_kernel(const float* __restrict__ U, int64_t ldu, int K, int R, int C,
                                                   double* __restrict__ G) {
    __shared__ double As[16][17], Bs[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int r = blockIdx.y * 16 + ty;     // row of G
    const int q = blockIdx.x * 16 + tx;     // column of G
    double acc = 0.0;
    for (int c0 = 0; c0 < C; c0 += 16) {
        const int ra = blockIdx.y * 16 + ty, rb = blockIdx.x * 16 + ty;
        As[ty][tx] = (ra < R && c0 + tx < C) ? (double)U[(int64_t)(ra / K) * ldu + (ra % K) * C + c0 + tx] : 0.0;
        Bs[ty][tx] = (rb < R && c0 + tx < C) ? (double)U[(int64_t)(rb / K) * ldu + (rb % K) * C + c0 + tx] : 0.0;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) acc += As[ty][e] * Bs[tx][e];
        __syncthreads();
    }
    if (r < R && q < R) G[(int64_t)r * R + q] = acc;
}

// ctc[j][l][i]: one thread per (i, j), all lags.  G is symmetric, so G[(i,k)][(j,m)] is read as
// G[(j,m)][(i,k)]: neighbouring threads (consecutive i) then read neighbouring K-runs of one row.
// K is a template parameter so that the K*K Gram entries of a pair stay in registers.
template <int KK>
__global__ __launch_bounds__(256) void ctc_kernel(const double* __restrict__ G, const double* __restrict__ wtw,
                                                  int T, int L, float* __restrict__ ctc, int64_t ldT) {
    extern __shared__ double swtw[];   // [K][K][L]
    for (int e = threadIdx.x; e < KK * KK * L; e += 256) swtw[e] = wtw[e];
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= T) return;
    const int R = T * KK;
    double g[KK * KK];
#pragma unroll
    for (int k = 0; k < KK; ++k)
#pragma unroll
        for (int m = 0; m < KK; ++m) g[k * KK + m] = G[(int64_t)(j * KK + m) * R + (i * KK + k)];
    for (int l = 0; l < L; ++l) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < KK * KK; ++q) acc += g[q] * swtw[q * L + l];
        ctc[((int64_t)j * L + l) * ldT + i] = (float)acc;
    }
}

template <int KK>
static hipError_t ctc_launch(const double* G, const double* wtw, int T, int L, float* ctc, int64_t ldT, hipStream_t s) {
    const size_t lds = sizeof(double) * KK * KK * L;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)ctc_kernel<KK>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((ctc_kernel<KK>), dim3((T + 255) / 256, T), dim3(256), lds, s, G, wtw, T, L, ctc, ldT);
    return hipGetLastError();
}

hipError_t launch_waves(const float* U, int64_t ldu, const float* wpca, int T, int K, int C, int nt,
                        float* waves, int64_t ldC, hipStream_t s) {
    const int64_t nw = (int64_t)T * nt * C;
    hipLaunchKernelGGL(waves_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, U, ldu, wpca, T, K, C, nt,
                       waves, ldC);
    return hipGetLastError();
}

// Features tables of prepare_extract (template_matching.py:20-26), one workgroup per template:
// iU[t] = argmax_c sum_k U[t][k][c]^2 (squares rounded, then summed in k order; first channel on
// ties) and Ucc[j][t][k] = U[t][k][iCC[j][iU[t]]] (iCC [nC][C]: channel iU's nearest channels).
__global__ __launch_bounds__(256) void extract_kernel(const float* __restrict__ U, int64_t ldu, int T, int K, int C,
                                                      const int* __restrict__ iCC, int nC, int* __restrict__ iU,
                                                      float* __restrict__ Ucc) {
    __shared__ float sv[256];
    __shared__ int si[256];
    const int t = blockIdx.x, tid = threadIdx.x;
    const float* u = U + (int64_t)t * ldu;
    float best = -1.f;
    int bi = 0x7fffffff;
    for (int c = tid; c < C; c += 256) {
        float e = 0.f;
        for (int k = 0; k < K; ++k) {
            const float x = u[(int64_t)k * C + c];
            e = __fadd_rn(e, __fmul_rn(x, x));
        }
        if (e > best) {   // ascending c within the thread: the first maximum stays
            best = e;
            bi = c;
        }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            const float v2 = sv[tid + off];
            const int i2 = si[tid + off];
            if (v2 > sv[tid] || (v2 == sv[tid] && i2 < si[tid])) {
                sv[tid] = v2;
                si[tid] = i2;
            }
        }
        __syncthreads();
    }
    const int b = (si[0] < C) ? si[0] : 0;   // no finite energy at all: channel 0
    if (tid == 0) iU[t] = b;
    for (int e2 = tid; e2 < nC * K; e2 += 256) {
        const int j = e2 / K, k = e2 - j * K;
        Ucc[((int64_t)j * T + t) * K + k] = u[(int64_t)k * C + iCC[(int64_t)j * C + b]];
    }
}

hipError_t launch_extract(const float* U, int64_t ldu, int T, int K, int C, const int* iCC, int nC, int* iU,
                          float* Ucc, hipStream_t s) {
    hipLaunchKernelGGL(extract_kernel, dim3(T), dim3(256), 0, s, U, ldu, T, K, C, iCC, nC, iU, Ucc);
    return hipGetLastError();
}

// Gram matrix, WtW, and the split-K partials of the Gram GEMM
size_t tables_workspace_doubles(int T, int K, int C, int nt) {
    const int R = T * K;
    int S;
    int64_t kchunk;
    gemm64_split(R, R, C, 1, &S, &kchunk);
    return (size_t)(1 + S) * R * R + (size_t)K * K * (2 * nt + 1);
}

// WtW[k][m][l] = sum_tau wPCA[k][nt - l + tau] wPCA[m][tau] (in range), the temporal part of ctc
// (prepare_matching, template_matching.py:110-121), in FP64 from the FP32 wPCA: each product of two
// widened FP32 values is exact, so the sum in tau order rounds as the host loop it replaces.
__global__ __launch_bounds__(256) void wtw_kernel(const float* __restrict__ wpca, int K, int nt,
                                                  double* __restrict__ wtw) {
    const int L = 2 * nt + 1;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= K * K * L) return;
    const int l = i % L, km = i / L, m = km % K, k = km / K;
    double acc = 0.0;
    for (int tau = 0; tau < nt; ++tau) {
        const int a = nt - l + tau;
        if (a >= 0 && a < nt) acc += (double)wpca[k * nt + a] * (double)wpca[m * nt + tau];
    }
    wtw[i] = acc;
}

hipError_t launch_tables(const float* U, int64_t ldu, const float* wpca, int T, int K, int C, int nt,
                         float* nm, float* waves, int64_t ldC, float* ctc, int64_t ldT, double* ws, hipStream_t s) {
    if (K > 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(nm_kernel, dim3(T), dim3(256), 0, s, U, ldu, K * C, nm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t nw = (int64_t)T * nt * C;
    hipLaunchKernelGGL(waves_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, U, ldu, wpca, T, K, C, nt, waves, ldC);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // WtW on the device (K*K*(2nt+1) values) from the FP32 wPCA
    const int L = 2 * nt + 1;
    const int R = T * K;
    double* dG = ws;
    double* dW = ws + (size_t)R * R;
    double* dP = dW + (size_t)K * K * L;
    hipLaunchKernelGGL(wtw_kernel, dim3((K * K * L + 255) / 256), dim3(256), 0, s, wpca, K, nt, dW);
    e = hipGetLastError();
    if (e == hipSuccess) {
        int S;
        int64_t kchunk;
        gemm64_split(R, R, C, 1, &S, &kchunk);
        if (ldu == (int64_t)K * C) {
            // U rows are a plain [R][C] matrix: the symmetric FP64 GEMM (FP32 inputs widened exactly)
            e = launch_gemm64_f32_nk(U, C, U, C, dP, R, R, C, S, kchunk, 1, s);
            if (e == hipSuccess) e = launch_reduce64(dP, S, R, R, 1, 1.0, 0.0, 0.0, dG, R, s);
        } else {
            hipLaunchKernelGGL(gram_kernel, dim3((R + 15) / 16, (R + 15) / 16), dim3(256), 0, s, U, ldu, K, R, C, dG);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) {
        switch (K) {
            case 1: e = ctc_launch<1>(dG, dW, T, L, ctc, ldT, s); break;
            case 2: e = ctc_launch<2>(dG, dW, T, L, ctc, ldT, s); break;
            case 3: e = ctc_launch<3>(dG, dW, T, L, ctc, ldT, s); break;
            case 4: e = ctc_launch<4>(dG, dW, T, L, ctc, ldT, s); break;
            case 5: e = ctc_launch<5>(dG, dW, T, L, ctc, ldT, s); break;
            case 6: e = ctc_launch<6>(dG, dW, T, L, ctc, ldT, s); break;
            case 7: e = ctc_launch<7>(dG, dW, T, L, ctc, ldT, s); break;
            default: e = ctc_launch<8>(dG, dW, T, L, ctc, ldT, s); break;
        }
    }
    hipError_t e2 = hipStreamSynchronize(s);
    return e != hipSuccess ? e : e2;
}

}  // namespace lss
