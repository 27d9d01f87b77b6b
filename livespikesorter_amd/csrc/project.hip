// Temporal PCA projection of the whitened window (Kilosort4 template_matching.py:145:
// conv1d(X, wPCA, padding=nt//2)), time-major:
//   Xp[w][k*C + c] = sum_m X[w + m - nt/2][c] * wPCA[k][m],  X = 0 outside [0, W).
// A workgroup owns 64 channels (one per lane) x PW samples; the PW+nt-1 input rows are staged in
// LDS once.  Each wavefront produces PW/4 samples in groups of R = 8: for every 8 taps it reads 16
// inputs of its channel from LDS and does 8 x 8 x K FMAs against the (wavefront-uniform) wPCA
// coefficients, so the kernel is FMA-bound rather than LDS-bound.  PW = 128 for long windows;
// short windows (the streaming batches) use PW = 32 so that the grid still covers the device.
#include "lss_kernels.h"

namespace lss {

constexpr int kProjC = 64;    // channels per workgroup (one per lane)
constexpr int kR = 8;         // samples per register group
constexpr int kMaxNt = 127;
constexpr int kMaxK = 8;

template <int KK, int PW>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ x, int64_t ldx, int64_t W,
                                                      int C, const float* __restrict__ wpca, int nt,
                                                      float* __restrict__ xp, int64_t ldp) {
    extern __shared__ float xs[];   // [PW + nt - 1 + 2 kR][kProjC]: staged rows incl. unrolled overrun
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t w0 = (int64_t)blockIdx.x * PW;
    const int c0 = blockIdx.y * kProjC;
    const int h = nt / 2;
    const int rows = PW + nt - 1;
    const int nrows = rows + 2 * kR;
    const int c = c0 + lane;
    for (int r = wave; r < nrows; r += 4) {
        const int64_t w = w0 - h + r;
        xs[r * kProjC + lane] = (r < rows && w >= 0 && w < W && c < C) ? x[w * ldx + c] : 0.f;
    }
    __syncthreads();
    if (c >= C) return;
    for (int g = 0; g < PW / 4 / kR; ++g) {
        const int s0 = wave * (PW / 4) + g * kR;
        if (w0 + s0 >= W) break;
        float acc[kR][KK];
#pragma unroll
        for (int r = 0; r < kR; ++r)
#pragma unroll
            for (int k = 0; k < KK; ++k) acc[r][k] = 0.f;
        for (int m0 = 0; m0 < nt; m0 += kR) {
            float xv[2 * kR];
#pragma unroll
            for (int i = 0; i < 2 * kR; ++i) xv[i] = xs[(s0 + m0 + i) * kProjC + lane];
#pragma unroll
            for (int mm = 0; mm < kR; ++mm) {
                const int m = m0 + mm;
                if (m < nt) {
                    float wk[KK];
#pragma unroll
                    for (int k = 0; k < KK; ++k) wk[k] = wpca[k * nt + m];
#pragma unroll
                    for (int r = 0; r < kR; ++r)
#pragma unroll
                        for (int k = 0; k < KK; ++k) acc[r][k] = fmaf(xv[mm + r], wk[k], acc[r][k]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            const int64_t w = w0 + s0 + r;
            if (w < W)
#pragma unroll
                for (int k = 0; k < KK; ++k) xp[w * ldp + (int64_t)k * C + c] = acc[r][k];
        }
    }
}

template <int KV, int PW>
hipError_t project_launch(hipStream_t s, const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int nt,
                          float* xp, int64_t ldp) {
    const dim3 grid((unsigned)((W + PW - 1) / PW), (C + kProjC - 1) / kProjC);
    const size_t lds = sizeof(float) * (size_t)(PW + nt - 1 + 2 * kR) * kProjC;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)project_kernel<KV, PW>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((project_kernel<KV, PW>), grid, dim3(256), lds, s, x_wc, ldx, W, C, wpca, nt, xp, ldp);
    return hipGetLastError();
}

// Windows of fewer than this many samples use 32-sample workgroups.
constexpr int64_t kProjShortW = 16384;

hipError_t launch_project(const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int K,
                          int nt, float* xp, int64_t ldp, hipStream_t s) {
    if (K < 1 || K > kMaxK || nt < 1 || nt > kMaxNt) return hipErrorInvalidValue;
    const bool shortw = W < kProjShortW;
    switch (K) {
#define LSS_PK(KV)                                                                                   \
    case KV:                                                                                         \
        return shortw ? project_launch<KV, 32>(s, x_wc, ldx, W, C, wpca, nt, xp, ldp)              \
                      : project_launch<KV, 128>(s, x_wc, ldx, W, C, wpca, nt, xp, ldp);
        LSS_PK(1) LSS_PK(2) LSS_PK(3) LSS_PK(4) LSS_PK(5) LSS_PK(6) LSS_PK(7) LSS_PK(8)
#undef LSS_PK
    }
    return hipErrorInvalidValue;
}

}  // namespace lss
