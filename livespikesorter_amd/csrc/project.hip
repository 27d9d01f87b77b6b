// Temporal PCA projection of the whitened window (Kilosort4 template_matching.py:145:
// conv1d(X, wPCA, padding=nt//2)), time-major:
//   Xp[w][k*C + c] = sum_m X[w + m - nt/2][c] * wPCA[k][m],  X = 0 outside [0, W).
// A workgroup owns 64 channels x 64 samples; the 64+nt-1 input rows are staged in LDS once and each
// thread keeps the K outputs of its (sample, channel) pairs in registers.
#include "lss_kernels.h"

namespace lss {

constexpr int kProjW = 64;    // samples per workgroup
constexpr int kProjC = 64;    // channels per workgroup (one per lane)
constexpr int kMaxNt = 127;
constexpr int kMaxK = 8;

template <int KK>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ x, int64_t ldx, int64_t W, int C,
                                                      const float* __restrict__ wpca, int nt,
                                                      float* __restrict__ xp, int64_t ldp) {
    __shared__ float xs[(kProjW + kMaxNt - 1) * kProjC];
    __shared__ float wp[kMaxK * kMaxNt];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t w0 = (int64_t)blockIdx.x * kProjW;
    const int c0 = blockIdx.y * kProjC;
    const int h = nt / 2;
    const int rows = kProjW + nt - 1;
    for (int i = threadIdx.x; i < KK * nt; i += 256) wp[i] = wpca[i];
    for (int i = threadIdx.x; i < rows * kProjC; i += 256) {
        int r = i / kProjC, cc = i % kProjC;
        int64_t w = w0 - h + r;
        int c = c0 + cc;
        xs[i] = (w >= 0 && w < W && c < C) ? x[w * ldx + c] : 0.f;
    }
    __syncthreads();
    const int c = c0 + lane;
    if (c >= C) return;
    // each wave: 16 consecutive samples
    for (int s = wave * 16; s < wave * 16 + 16; ++s) {
        const int64_t w = w0 + s;
        if (w >= W) break;
        float acc[KK];
#pragma unroll
        for (int k = 0; k < KK; ++k) acc[k] = 0.f;
        for (int m = 0; m < nt; ++m) {
            const float v = xs[(s + m) * kProjC + lane];
#pragma unroll
            for (int k = 0; k < KK; ++k) acc[k] = fmaf(v, wp[k * nt + m], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < KK; ++k) xp[w * ldp + (int64_t)k * C + c] = acc[k];
    }
}

hipError_t launch_project(const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int K,
                          int nt, float* xp, int64_t ldp, hipStream_t s) {
    if (K < 1 || K > kMaxK || nt < 1 || nt > kMaxNt) return hipErrorInvalidValue;
    dim3 grid((unsigned)((W + kProjW - 1) / kProjW), (C + kProjC - 1) / kProjC);
    switch (K) {
#define LSS_PK(KV) case KV: hipLaunchKernelGGL(project_kernel<KV>, grid, dim3(256), 0, s, x_wc, ldx, W, C, wpca, nt, xp, ldp); break;
        LSS_PK(1) LSS_PK(2) LSS_PK(3) LSS_PK(4) LSS_PK(5) LSS_PK(6) LSS_PK(7) LSS_PK(8)
#undef LSS_PK
    }
    return hipGetLastError();
}

}  // namespace lss
