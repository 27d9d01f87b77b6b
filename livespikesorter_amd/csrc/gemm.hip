// Dense contractions of the window on CDNA4 matrix cores.
//
//   whitening : Xw[w][c'] = sum_c X[c][w] * Wm[c'][c]            (M = W, N = C, K = C)
//   scores    : B[w][t]   = sum_q Xp[w][q] * U[t][q], q = k*C+c   (M = W, N = T, K = K*C)
//
// both as C[m][n] = sum_k A(m,k) B(n,k).  A 128x128 output tile per 256-thread workgroup; each of
// the 4 wavefronts owns a 64x64 quarter as 2x2 blocks of 32x32 MFMA accumulators.  Operand tiles
// are staged through LDS k-major so that every MFMA operand read is 32 consecutive floats.
// Workgroups are renumbered so that each XCD (blockIdx % 8) walks a contiguous run of tiles with
// the N tiles of one M row-block adjacent: the A rows are then fetched into that XCD's L2 once.
#include "lss_kernels.h"

namespace lss {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 16;

__device__ __forceinline__ void xcd_tile(int nM, int nN, int& mt, int& nt) {
    const int total = nM * nN;
    const int b = blockIdx.x;
    const int per = (total + 7) / 8;
    const int xcd = b % 8, idx = b / 8;
    // the grid is 8*per blocks: v is a bijection onto [0, 8*per); v >= total is idle
    const int v = xcd * per + idx;
    mt = (v < total) ? v / nN : nM;
    nt = v % nN;
}

template <bool A_KM>
__global__ __launch_bounds__(256) void gemm_f32_kernel(const float* __restrict__ A, int64_t lda,
                                                       const float* __restrict__ B, int64_t ldb,
                                                       float* __restrict__ C, int64_t ldc,
                                                       int M, int N, int K, int nM, int nN) {
    __shared__ float As[BK][BM + 4];
    __shared__ float Bs[BK][BN + 4];
    int mt, ntile;
    xcd_tile(nM, nN, mt, ntile);
    if (mt >= nM) return;
    const int m0 = mt * BM, n0 = ntile * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    for (int k0 = 0; k0 < K; k0 += BK) {
        // ---- stage A ----
        if (!A_KM) {
            const int row = tid >> 1, kq = (tid & 1) * 8;
            const int m = m0 + row;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = k0 + kq + 4 * h;
                floatx4 v = {0.f, 0.f, 0.f, 0.f};
                if (m < M) {
                    if (k + 3 < K) v = *(const floatx4*)(A + (int64_t)m * lda + k);
                    else
                        for (int e = 0; e < 4; ++e) v[e] = (k + e < K) ? A[(int64_t)m * lda + k + e] : 0.f;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) As[kq + 4 * h + e][row] = v[e];
            }
        } else {
            const int kr = tid >> 4, mq = (tid & 15) * 8;
            const int k = k0 + kr;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int m = m0 + mq + 4 * h;
                floatx4 v = {0.f, 0.f, 0.f, 0.f};
                if (k < K) {
                    if (m + 3 < M) v = *(const floatx4*)(A + (int64_t)k * lda + m);
                    else
                        for (int e = 0; e < 4; ++e) v[e] = (m + e < M) ? A[(int64_t)k * lda + m + e] : 0.f;
                }
                *(floatx4*)&As[kr][mq + 4 * h] = v;
            }
        }
        // ---- stage B (N x K, k contiguous) ----
        {
            const int row = tid >> 1, kq = (tid & 1) * 8;
            const int n = n0 + row;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = k0 + kq + 4 * h;
                floatx4 v = {0.f, 0.f, 0.f, 0.f};
                if (n < N) {
                    if (k + 3 < K) v = *(const floatx4*)(B + (int64_t)n * ldb + k);
                    else
                        for (int e = 0; e < 4; ++e) v[e] = (k + e < K) ? B[(int64_t)n * ldb + k + e] : 0.f;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) Bs[kq + 4 * h + e][row] = v[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
            const int kr = kk + (lane >> 5);
            float a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kr][wm * 64 + i * 32 + (lane & 31)];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn * 64 + j * 32 + (lane & 31)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // ---- store: acc[r] of lane l is row 8*(r/4) + 4*(l/32) + r%4, column l%32 ----
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + wn * 64 + j * 32 + (lane & 31);
            if (n >= N) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m < M) C[(int64_t)m * ldc + n] = acc[i][j][r];
            }
        }
}

hipError_t launch_gemm_f32(const float* A, int64_t lda, int a_km, const float* B, int64_t ldb,
                           float* C, int64_t ldc, int M, int N, int K, hipStream_t s) {
    if ((lda & 3) || (ldb & 3)) return hipErrorInvalidValue;
    const int nM = (M + BM - 1) / BM, nN = (N + BN - 1) / BN;
    dim3 grid(8 * ((nM * nN + 7) / 8));
    if (a_km)
        hipLaunchKernelGGL(gemm_f32_kernel<true>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K, nM, nN);
    else
        hipLaunchKernelGGL(gemm_f32_kernel<false>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K, nM, nN);
    return hipGetLastError();
}

}  // namespace lss
