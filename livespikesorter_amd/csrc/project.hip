// Temporal PCA projection of the whitened window (Kilosort4 template_matching.py:145:
// conv1d(X, wPCA, padding=nt//2)), time-major:
//   Xp[w][k*C + c] = sum_m X[w + m - nt/2][c] * wPCA[k][m],  X = 0 outside [0, W).
// A workgroup owns 64 channels (one per lane) x PW samples; the PW+nt-1 input rows are staged in
// LDS once.  Each wavefront produces PW/4 samples in groups of R = 8: for every 8 taps it reads 16
// inputs of its channel from LDS and does 8 x 8 x K FMAs against the (wavefront-uniform) wPCA
// coefficients, so the kernel is FMA-bound rather than LDS-bound.  PW = 128 for long windows;
// short windows (the streaming batches) use PW = 32 so that the grid still covers the device.
#include "../../include/lss.h"
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

// tile: 0 = by window length (32 samples below kProjShortW, else 128), 1 = 32, 2 = 64, 3 = 128
// samples per workgroup (every output is the same FMA chain over m whatever the shape)
hipError_t launch_project(const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int K,
                          int nt, float* xp, int64_t ldp, hipStream_t s, int tile) {
    if (K < 1 || K > kMaxK || nt < 1 || nt > kMaxNt || tile < 0 || tile > 3) return hipErrorInvalidValue;
    const int pw = tile == 1 ? 32 : tile == 2 ? 64 : tile == 3 ? 128 : (W < kProjShortW ? 32 : 128);
    switch (K) {
#define LSS_PK(KV)                                                                                   \
    case KV:                                                                                         \
        return pw == 32 ? project_launch<KV, 32>(s, x_wc, ldx, W, C, wpca, nt, xp, ldp)            \
             : pw == 64 ? project_launch<KV, 64>(s, x_wc, ldx, W, C, wpca, nt, xp, ldp)            \
                        : project_launch<KV, 128>(s, x_wc, ldx, W, C, wpca, nt, xp, ldp);
        LSS_PK(1) LSS_PK(2) LSS_PK(3) LSS_PK(4) LSS_PK(5) LSS_PK(6) LSS_PK(7) LSS_PK(8)
#undef LSS_PK
    }
    return hipErrorInvalidValue;
}

// ---- split-precision projection on the matrix cores ---------------------------------------------
// For the bf16x3 / fp16x3 paths.  A workgroup covers 32 output samples x 64 channels with one
// wavefront per PC k (64*K threads).  Per k the 32 x 128 output block is a banded product
//   Xp[w0+s][k*C+c] = sum_j T_k[s][j] Xb[j][c],   T_k[s][j] = wPCA[k][j-s] (0 <= j-s < nt),
//   Xb[j][c] = X[w0 - nt/2 + j][c] (zero outside [0, W)),   j < J = roundup(32 + nt - 1, 16),
// on 32x32x16 MFMA with the split operands x = hi + lo and hi*hi + hi*lo + lo*hi accumulated in FP32
// (the split of the score GEMM).  Each lane forms its A fragment (8 band entries) from wPCA in LDS
// once per K step and uses it for both 32-channel blocks; X is staged transposed (channel rows,
// j contiguous) as hi/lo halves.
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 pbf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 phalfx8 __attribute__((ext_vector_type(8)));

constexpr int kPxS = 32;      // output samples per workgroup
// NB 32-channel blocks per workgroup: 2 by default (64 channels, 28 KB of LDS, so more workgroups
// stay resident than with 128 on 2-s windows: 0.32 -> 0.24 ms); 1 and 4 for the tile autotuner

template <int PREC>
struct PxType;
template <>
struct PxType<LSS_PREC_BF16X3> { typedef __bf16 T; typedef pbf16x8 V; };
template <>
struct PxType<LSS_PREC_FP16X3> { typedef _Float16 T; typedef phalfx8 V; };

template <int PREC>
__device__ __forceinline__ floatx16 px_mfma(typename PxType<PREC>::V a, typename PxType<PREC>::V b, floatx16 c) {
    if constexpr (PREC == LSS_PREC_FP16X3) return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

template <int PREC, int NB>
__global__ __launch_bounds__(512) void project_x3_kernel(const float* __restrict__ x, int64_t ldx, int64_t W, int C,
                                                         const float* __restrict__ wpca, int K, int nt,
                                                         float* __restrict__ xp, int64_t ldp, int J) {
    typedef typename PxType<PREC>::T T;
    typedef typename PxType<PREC>::V V;
    constexpr int kPxNB = NB, kPxC = 32 * NB;
    extern __shared__ __attribute__((aligned(16))) unsigned char px_smem[];
    const int ldj = J + 8;                             // 16-byte aligned rows
    T* Xh = (T*)px_smem;                               // [kPxC][ldj]
    T* Xl = Xh + kPxC * ldj;
    float* wp = (float*)(Xl + kPxC * ldj);             // [K][nt]
    const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63, k = tid >> 6;
    const int64_t w0 = (int64_t)blockIdx.x * kPxS;
    const int c0 = blockIdx.y * kPxC;
    const int h = nt / 2;
    for (int e = tid; e < K * nt; e += nth) wp[e] = wpca[e];
    for (int e = tid; e < kPxC * J; e += nth) {          // consecutive lanes read consecutive channels
        const int cc = e % kPxC, jj = e / kPxC;
        const int c = c0 + cc;
        {
            const int64_t w = w0 - h + jj;
            const float v = (w >= 0 && w < W && c < C) ? x[w * ldx + c] : 0.f;
            const T hi = (T)v;
            Xh[cc * ldj + jj] = hi;
            Xl[cc * ldj + jj] = (T)(v - (float)hi);
        }
    }
    __syncthreads();
    const int m = lane & 31, g = lane >> 5;
    floatx16 acc[kPxNB];
#pragma unroll
    for (int b = 0; b < kPxNB; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[b][r] = 0.f;
    const float* wk = wp + k * nt;
    for (int kk = 0; kk < J; kk += 16) {
        V ah, al;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int d = kk + 8 * g + e - m;                // band entry T_k[m][j], j = kk + 8g + e
            const float a = (d >= 0 && d < nt) ? wk[d] : 0.f;
            const T hi = (T)a;
            ah[e] = hi;
            al[e] = (T)(a - (float)hi);
        }
#pragma unroll
        for (int b = 0; b < kPxNB; ++b) {
            const int row = (b * 32 + m) * ldj + kk + 8 * g;
            const V bh = *(const V*)&Xh[row];
            const V bl = *(const V*)&Xl[row];
            acc[b] = px_mfma<PREC>(al, bh, acc[b]);
            acc[b] = px_mfma<PREC>(ah, bl, acc[b]);
            acc[b] = px_mfma<PREC>(ah, bh, acc[b]);
        }
    }
    // acc[b][r] of lane l: sample 8*(r/4) + 4*(l/32) + r%4, channel c0 + 32b + l%32
#pragma unroll
    for (int b = 0; b < kPxNB; ++b) {
        const int c = c0 + b * 32 + m;
        if (c >= C) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t w = w0 + 8 * (r >> 2) + 4 * g + (r & 3);
            if (w < W) xp[w * ldp + (int64_t)k * C + c] = acc[b][r];
        }
    }
}

template <int PREC, int NB>
static hipError_t project_x3_launch(const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int K, int nt,
                                    float* xp, int64_t ldp, hipStream_t s) {
    constexpr int kPxC = 32 * NB;
    const int J = (kPxS + nt - 1 + 15) / 16 * 16;
    const size_t lds = 2 * sizeof(typename PxType<PREC>::T) * (size_t)kPxC * (J + 8) + sizeof(float) * K * nt;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)project_x3_kernel<PREC, NB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const dim3 grid((unsigned)((W + kPxS - 1) / kPxS), (C + kPxC - 1) / kPxC);
    hipLaunchKernelGGL((project_x3_kernel<PREC, NB>), grid, dim3(64 * K), lds, s, x_wc, ldx, W, C, wpca, K, nt, xp,
                       ldp, J);
    return hipGetLastError();
}

template <int PREC>
static hipError_t project_x3_tile(const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int K, int nt,
                                  float* xp, int64_t ldp, hipStream_t s, int tile) {
    switch (tile) {
        case 0:
        case 2: return project_x3_launch<PREC, 2>(x_wc, ldx, W, C, wpca, K, nt, xp, ldp, s);
        case 1: return project_x3_launch<PREC, 1>(x_wc, ldx, W, C, wpca, K, nt, xp, ldp, s);
        case 3: return project_x3_launch<PREC, 4>(x_wc, ldx, W, C, wpca, K, nt, xp, ldp, s);
        default: return hipErrorInvalidValue;
    }
}

// tile: 0 or 2 = 64 channels per workgroup, 1 = 32, 3 = 128 (per-element sums identical)
hipError_t launch_project_x3(int prec, const float* x_wc, int64_t ldx, int64_t W, int C, const float* wpca, int K,
                             int nt, float* xp, int64_t ldp, hipStream_t s, int tile) {
    if (K < 1 || K > kMaxK || nt < 1 || nt > kMaxNt) return hipErrorInvalidValue;
    if (prec == LSS_PREC_BF16X3)
        return project_x3_tile<LSS_PREC_BF16X3>(x_wc, ldx, W, C, wpca, K, nt, xp, ldp, s, tile);
    if (prec == LSS_PREC_FP16X3)
        return project_x3_tile<LSS_PREC_FP16X3>(x_wc, ldx, W, C, wpca, K, nt, xp, ldp, s, tile);
    return hipErrorInvalidValue;
}

}  // namespace lss
