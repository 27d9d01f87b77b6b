// Dense contractions of the window on CDNA4 matrix cores.
//
//   whitening : Xw[w][c'] = sum_c X[c][w] * Wm[c'][c]            (M = W, N = C, K = C)
//   scores    : B[w][t]   = sum_q Xp[w][q] * U[t][q], q = k*C+c   (M = W, N = T, K = K*C)
//
// both as C[m][n] = sum_k A(m,k) B(n,k).  A 128x128 output tile per 256-thread workgroup; each of
// the 4 wavefronts owns a 64x64 quarter as 2x2 blocks of 32x32 MFMA accumulators.  The K loop
// advances 32 at a time with register prefetch: the global loads of step k+1 are issued before
// the MFMAs of step k, so load latency hides behind matrix work.  Workgroups are renumbered so
// that each XCD (blockIdx % 8) walks a contiguous run of tiles with the N tiles of one M
// row-block adjacent: the A rows are fetched into that XCD's L2 once.
//
// FP32 uses v_mfma_f32_32x32x2_f32 on k-major LDS tiles.  The reduced-precision variants convert
// the FP32 operands while staging them into LDS (m/n-major, k contiguous, 8 values per lane read)
// and use v_mfma_f32_32x32x16_{bf16,f16}; bf16x3 splits x = hi + lo (hi = bf16(x),
// lo = bf16(x - hi)) and accumulates hi*hi + hi*lo + lo*hi in FP32; fp16x3 is the same split with
// FP16 halves (lo below FP16's normal range loses bits: its error is reported, not assumed).
#include "../../include/lss.h"
#include "lss_kernels.h"

namespace lss {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128, BK = 32;

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

// ---- global tile fetch: 128 rows x 32 k of an MK (row = m, k contiguous) or KM operand -------
// MK: thread t loads rows t/2 (+64 for the second half), k quarter (t&1)*16 .. +16 as 4 float4.
// KM: thread t loads k row t/8, m quarter (t&7)*16 .. +16 as 4 float4.
struct Frag {
    floatx4 v[4];
};

template <bool KM>
__device__ __forceinline__ void fetch(const float* __restrict__ P, int64_t ld, int r0, int k0, int R, int K,
                                      Frag& f) {
    const int tid = threadIdx.x;
    if (!KM) {
        const int row = tid >> 1, kq = (tid & 1) * 16;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int r = r0 + row;
            const int k = k0 + kq + 4 * h;
            floatx4 v = {0.f, 0.f, 0.f, 0.f};
            if (r < R) {
                const float* p = P + (int64_t)r * ld + k;
                if (k + 3 < K) v = *(const floatx4*)p;
                else
                    for (int e = 0; e < 4; ++e) v[e] = (k + e < K) ? p[e] : 0.f;
            }
            f.v[h] = v;
        }
        (void)R;
    } else {
        const int kr = tid >> 3, mq = (tid & 7) * 16;
        const int k = k0 + kr;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int r = r0 + mq + 4 * h;
            floatx4 v = {0.f, 0.f, 0.f, 0.f};
            if (k < K) {
                const float* p = P + (int64_t)k * ld + r;
                if (r + 3 < R) v = *(const floatx4*)p;
                else
                    for (int e = 0; e < 4; ++e) v[e] = (r + e < R) ? p[e] : 0.f;
            }
            f.v[h] = v;
        }
    }
}

// FP32: LDS tile [BK][128 + 4] (k-major)
template <bool KM>
__device__ __forceinline__ void stash_f32(float (*S)[BM + 4], const Frag& f) {
    const int tid = threadIdx.x;
    if (!KM) {
        const int row = tid >> 1, kq = (tid & 1) * 16;
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) S[kq + 4 * h + e][row] = f.v[h][e];
    } else {
        const int kr = tid >> 3, mq = (tid & 7) * 16;
#pragma unroll
        for (int h = 0; h < 4; ++h) *(floatx4*)&S[kr][mq + 4 * h] = f.v[h];
    }
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
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    Frag fa, fb;
    fetch<A_KM>(A, lda, m0, 0, M, K, fa);
    fetch<false>(B, ldb, n0, 0, N, K, fb);
    for (int k0 = 0; k0 < K; k0 += BK) {
        stash_f32<A_KM>(As, fa);
        stash_f32<false>(Bs, fb);
        __syncthreads();
        if (k0 + BK < K) {
            fetch<A_KM>(A, lda, m0, k0 + BK, M, K, fa);
            fetch<false>(B, ldb, n0, k0 + BK, N, K, fb);
        }
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
    // acc[r] of lane l is row 8*(r/4) + 4*(l/32) + r%4, column l%32 of its 32x32 block
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

// ---- reduced precision ------------------------------------------------------------------------
constexpr int kLdsK = BK + 8;   // k stride of the m/n-major LDS tiles (16-bit elements)

template <int PREC>
struct LpType;
template <>
struct LpType<LSS_PREC_BF16> { typedef __bf16 T; typedef bf16x8 V; };
template <>
struct LpType<LSS_PREC_FP16> { typedef _Float16 T; typedef halfx8 V; };
template <>
struct LpType<LSS_PREC_BF16X3> { typedef __bf16 T; typedef bf16x8 V; };
template <>
struct LpType<LSS_PREC_FP16X3> { typedef _Float16 T; typedef halfx8 V; };

template <int PREC>
constexpr bool is_split() { return PREC == LSS_PREC_BF16X3 || PREC == LSS_PREC_FP16X3; }

// Global fetch of the reduced-precision kernel: a k-major (KM) operand as fetch<true>; an m/n-major
// one with eight lanes per row, so that one load instruction covers 8 rows' 128-byte k runs (8
// cache lines) instead of 32 rows' halves: thread t loads rows t/8 + 32h, k (t&7)*4 .. +4
// (bf16x3 scores 0.68 -> 0.61 ms, plain BF16 0.56 -> 0.42 ms).  The k-major operand kept its
// layout (32 lanes per k row made the whitening GEMM slower, 0.11 -> 0.17 ms), and so did the
// FP32 kernel (eight lanes per row: 1.71 vs 1.67 ms; its MFMA work is 16x slower per element).
template <bool KM>
__device__ __forceinline__ void fetch_lp(const float* __restrict__ P, int64_t ld, int r0, int k0, int R, int K,
                                         Frag& f) {
    if (KM) {
        fetch<true>(P, ld, r0, k0, R, K, f);
        return;
    }
    const int tid = threadIdx.x;
    const int row = tid >> 3, k = k0 + (tid & 7) * 4;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const int r = r0 + row + 32 * h;
        floatx4 v = {0.f, 0.f, 0.f, 0.f};
        if (r < R) {
            const float* p = P + (int64_t)r * ld + k;
            if (k + 3 < K) v = *(const floatx4*)p;
            else
                for (int e = 0; e < 4; ++e) v[e] = (k + e < K) ? p[e] : 0.f;
        }
        f.v[h] = v;
    }
}

// element (row, k) of the tile held by this thread's Frag (as fetch_lp loads it)
template <bool KM>
__device__ __forceinline__ void frag_coord(int h, int e, int& row, int& k) {
    const int tid = threadIdx.x;
    if (!KM) {
        row = (tid >> 3) + 32 * h;
        k = (tid & 7) * 4 + e;
    } else {
        row = (tid & 7) * 16 + 4 * h + e;
        k = tid >> 3;
    }
}

template <int PREC, bool KM>
__device__ __forceinline__ void stash_lp(typename LpType<PREC>::T (*Hi)[kLdsK], typename LpType<PREC>::T (*Lo)[kLdsK],
                                         const Frag& f) {
    typedef typename LpType<PREC>::T T;
#pragma unroll
    for (int h = 0; h < 4; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int row, k;
            frag_coord<KM>(h, e, row, k);
            const float x = f.v[h][e];
            const T hi = (T)x;
            Hi[row][k] = hi;
            if (is_split<PREC>()) Lo[row][k] = (T)(x - (float)hi);
        }
}

template <int PREC>
__device__ __forceinline__ floatx16 mfma_lp(typename LpType<PREC>::V a, typename LpType<PREC>::V b, floatx16 c) {
    if constexpr (PREC == LSS_PREC_FP16 || PREC == LSS_PREC_FP16X3)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Three workgroups per CU (registers capped at 168; the 40 KB of LDS fits three times): the
// extra wavefront per SIMD hides more of the tile loads (bf16x3 scores 0.71 -> 0.68 ms).
template <int PREC, bool A_KM>
__global__ __launch_bounds__(256, 3) void gemm_lp_kernel(const float* __restrict__ A, int64_t lda,
                                                      const float* __restrict__ B, int64_t ldb,
                                                      float* __restrict__ C, int64_t ldc,
                                                      int M, int N, int K, int nM, int nN) {
    typedef typename LpType<PREC>::T T;
    typedef typename LpType<PREC>::V V;
    constexpr bool X3 = is_split<PREC>();
    __shared__ __attribute__((aligned(16))) T Ah[BM][kLdsK];
    __shared__ __attribute__((aligned(16))) T Bh[BN][kLdsK];
    __shared__ __attribute__((aligned(16))) T Al[X3 ? BM : 1][kLdsK];
    __shared__ __attribute__((aligned(16))) T Bl[X3 ? BN : 1][kLdsK];
    int mt, ntile;
    xcd_tile(nM, nN, mt, ntile);
    if (mt >= nM) return;
    const int m0 = mt * BM, n0 = ntile * BN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    Frag fa, fb;
    fetch_lp<A_KM>(A, lda, m0, 0, M, K, fa);
    fetch_lp<false>(B, ldb, n0, 0, N, K, fb);
    for (int k0 = 0; k0 < K; k0 += BK) {
        stash_lp<PREC, A_KM>(Ah, Al, fa);
        stash_lp<PREC, false>(Bh, Bl, fb);
        __syncthreads();
        if (k0 + BK < K) {
            fetch_lp<A_KM>(A, lda, m0, k0 + BK, M, K, fa);
            fetch_lp<false>(B, ldb, n0, k0 + BK, N, K, fb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int kq = kk + 8 * (lane >> 5);
            V ah[2], bh[2], al[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int r = wm * 64 + i * 32 + (lane & 31);
                ah[i] = *(const V*)&Ah[r][kq];
                if constexpr (X3) al[i] = *(const V*)&Al[r][kq];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int r = wn * 64 + j * 32 + (lane & 31);
                bh[j] = *(const V*)&Bh[r][kq];
                if constexpr (X3) bl[j] = *(const V*)&Bl[r][kq];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if constexpr (X3) {
                        acc[i][j] = mfma_lp<PREC>(al[i], bh[j], acc[i][j]);
                        acc[i][j] = mfma_lp<PREC>(ah[i], bl[j], acc[i][j]);
                    }
                    acc[i][j] = mfma_lp<PREC>(ah[i], bh[j], acc[i][j]);
                }
        }
        __syncthreads();
    }
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

template <int PREC>
hipError_t lp_launch(const float* A, int64_t lda, int a_km, const float* B, int64_t ldb, float* C, int64_t ldc,
                     int M, int N, int K, hipStream_t s) {
    const int nM = (M + BM - 1) / BM, nN = (N + BN - 1) / BN;
    dim3 grid(8 * ((nM * nN + 7) / 8));
    if (a_km)
        hipLaunchKernelGGL((gemm_lp_kernel<PREC, true>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K, nM, nN);
    else
        hipLaunchKernelGGL((gemm_lp_kernel<PREC, false>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K, nM, nN);
    return hipGetLastError();
}

hipError_t launch_gemm_lp(int prec, const float* A, int64_t lda, int a_km, const float* B, int64_t ldb,
                          float* C, int64_t ldc, int M, int N, int K, hipStream_t s) {
    if ((lda & 3) || (ldb & 3)) return hipErrorInvalidValue;
    switch (prec) {
        case LSS_PREC_BF16: return lp_launch<LSS_PREC_BF16>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case LSS_PREC_FP16: return lp_launch<LSS_PREC_FP16>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case LSS_PREC_BF16X3: return lp_launch<LSS_PREC_BF16X3>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case LSS_PREC_FP16X3: return lp_launch<LSS_PREC_FP16X3>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lss
