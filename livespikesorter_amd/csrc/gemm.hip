// Dense contractions of the window on CDNA4 matrix cores.
//
//   whitening : Xw[w][c'] = sum_c X[c][w] * Wm[c'][c]            (M = W, N = C, K = C)
//   scores    : B[w][t]   = sum_q Xp[w][q] * U[t][q], q = k*C+c   (M = W, N = T, K = K*C)
//
// both as C[m][n] = sum_k A(m,k) B(n,k).  A TM x TN output tile per 256-thread workgroup (TM, TN in
// {64, 128}: four shapes, picked per handle by the tile autotuner); each of the 4 wavefronts owns a
// (TM/2) x (TN/2) quarter as blocks of 32x32 MFMA accumulators.  Every output element belongs to
// one 32x32 block whose K loop runs in the same order for every shape, so all shapes give
// bit-identical results; small tiles fill the device on short windows.  The K loop
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

constexpr int BK = 32;
constexpr int kMaxTile = 128;

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

// ---- global tile fetch: ROWS rows (64 or 128) x 32 k of an MK (row = m, k contiguous) or KM operand
// MK: thread t loads row t/(256/ROWS), k range (t % (256/ROWS)) * (ROWS/8) .. + ROWS/8 as float4s.
// KM: thread t loads k row t/8, m range (t&7) * (ROWS/8) .. + ROWS/8 as float4s.
struct Frag {
    floatx4 v[4];
};

template <bool KM, int ROWS>
__device__ __forceinline__ void fetch(const float* __restrict__ P, int64_t ld, int r0, int k0, int R, int K,
                                      Frag& f) {
    constexpr int H = ROWS / 32;             // float4 per thread
    const int tid = threadIdx.x;
    if (!KM) {
        constexpr int TPR = 256 / ROWS;      // threads per row
        const int row = tid / TPR, kq = (tid % TPR) * (4 * H);
#pragma unroll
        for (int h = 0; h < H; ++h) {
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
    } else {
        const int kr = tid >> 3, mq = (tid & 7) * (4 * H);
        const int k = k0 + kr;
#pragma unroll
        for (int h = 0; h < H; ++h) {
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

// FP32: LDS tile [BK][ROWS + 4] (k-major)
template <bool KM, int ROWS>
__device__ __forceinline__ void stash_f32(float (*S)[kMaxTile + 4], const Frag& f) {
    constexpr int H = ROWS / 32;
    const int tid = threadIdx.x;
    if (!KM) {
        constexpr int TPR = 256 / ROWS;
        const int row = tid / TPR, kq = (tid % TPR) * (4 * H);
#pragma unroll
        for (int h = 0; h < H; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) S[kq + 4 * h + e][row] = f.v[h][e];
    } else {
        const int kr = tid >> 3, mq = (tid & 7) * (4 * H);
#pragma unroll
        for (int h = 0; h < H; ++h) *(floatx4*)&S[kr][mq + 4 * h] = f.v[h];
    }
}

// TM x TN tile: wavefront (wm, wn) owns rows wm*TM/2 .. as I = TM/64 blocks of 32 and columns
// wn*TN/2 .. as J = TN/64 blocks
template <bool A_KM, int TM, int TN>
__global__ __launch_bounds__(256) void gemm_f32_kernel(const float* __restrict__ A, int64_t lda,
                                                       const float* __restrict__ B, int64_t ldb,
                                                       float* __restrict__ C, int64_t ldc,
                                                       int M, int N, int K, int nM, int nN) {
    constexpr int I = TM / 64, J = TN / 64;
    __shared__ float As[BK][kMaxTile + 4];
    __shared__ float Bs[BK][kMaxTile + 4];
    int mt, ntile;
    xcd_tile(nM, nN, mt, ntile);
    if (mt >= nM) return;
    const int m0 = mt * TM, n0 = ntile * TN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    floatx16 acc[I][J];
#pragma unroll
    for (int i = 0; i < I; ++i)
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    Frag fa, fb;
    fetch<A_KM, TM>(A, lda, m0, 0, M, K, fa);
    fetch<false, TN>(B, ldb, n0, 0, N, K, fb);
    for (int k0 = 0; k0 < K; k0 += BK) {
        stash_f32<A_KM, TM>(As, fa);
        stash_f32<false, TN>(Bs, fb);
        __syncthreads();
        if (k0 + BK < K) {
            fetch<A_KM, TM>(A, lda, m0, k0 + BK, M, K, fa);
            fetch<false, TN>(B, ldb, n0, k0 + BK, N, K, fb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
            const int kr = kk + (lane >> 5);
            float a[I], b[J];
#pragma unroll
            for (int i = 0; i < I; ++i) a[i] = As[kr][wm * (TM / 2) + i * 32 + (lane & 31)];
#pragma unroll
            for (int j = 0; j < J; ++j) b[j] = Bs[kr][wn * (TN / 2) + j * 32 + (lane & 31)];
#pragma unroll
            for (int i = 0; i < I; ++i)
#pragma unroll
                for (int j = 0; j < J; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // acc[r] of lane l is row 8*(r/4) + 4*(l/32) + r%4, column l%32 of its 32x32 block
#pragma unroll
    for (int i = 0; i < I; ++i)
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int n = n0 + wn * (TN / 2) + j * 32 + (lane & 31);
            if (n >= N) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * (TM / 2) + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
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
template <bool KM, int ROWS>
__device__ __forceinline__ void fetch_lp(const float* __restrict__ P, int64_t ld, int r0, int k0, int R, int K,
                                         Frag& f) {
    if (KM) {
        fetch<true, ROWS>(P, ld, r0, k0, R, K, f);
        return;
    }
    const int tid = threadIdx.x;
    const int row = tid >> 3, k = k0 + (tid & 7) * 4;
#pragma unroll
    for (int h = 0; h < ROWS / 32; ++h) {
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
template <bool KM, int ROWS>
__device__ __forceinline__ void frag_coord(int h, int e, int& row, int& k) {
    const int tid = threadIdx.x;
    if (!KM) {
        row = (tid >> 3) + 32 * h;
        k = (tid & 7) * 4 + e;
    } else {
        row = (tid & 7) * (ROWS / 8) + 4 * h + e;
        k = tid >> 3;
    }
}

template <int PREC, bool KM, int ROWS>
__device__ __forceinline__ void stash_lp(typename LpType<PREC>::T (*Hi)[kLdsK], typename LpType<PREC>::T (*Lo)[kLdsK],
                                         const Frag& f) {
    typedef typename LpType<PREC>::T T;
#pragma unroll
    for (int h = 0; h < ROWS / 32; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int row, k;
            frag_coord<KM, ROWS>(h, e, row, k);
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
template <int PREC, bool A_KM, int TM, int TN>
__global__ __launch_bounds__(256, 3) void gemm_lp_kernel(const float* __restrict__ A, int64_t lda,
                                                      const float* __restrict__ B, int64_t ldb,
                                                      float* __restrict__ C, int64_t ldc,
                                                      int M, int N, int K, int nM, int nN) {
    typedef typename LpType<PREC>::T T;
    typedef typename LpType<PREC>::V V;
    constexpr bool X3 = is_split<PREC>();
    constexpr int I = TM / 64, J = TN / 64;
    __shared__ __attribute__((aligned(16))) T Ah[TM][kLdsK];
    __shared__ __attribute__((aligned(16))) T Bh[TN][kLdsK];
    __shared__ __attribute__((aligned(16))) T Al[X3 ? TM : 1][kLdsK];
    __shared__ __attribute__((aligned(16))) T Bl[X3 ? TN : 1][kLdsK];
    int mt, ntile;
    xcd_tile(nM, nN, mt, ntile);
    if (mt >= nM) return;
    const int m0 = mt * TM, n0 = ntile * TN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    floatx16 acc[I][J];
#pragma unroll
    for (int i = 0; i < I; ++i)
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    Frag fa, fb;
    fetch_lp<A_KM, TM>(A, lda, m0, 0, M, K, fa);
    fetch_lp<false, TN>(B, ldb, n0, 0, N, K, fb);
    for (int k0 = 0; k0 < K; k0 += BK) {
        stash_lp<PREC, A_KM, TM>(Ah, Al, fa);
        stash_lp<PREC, false, TN>(Bh, Bl, fb);
        __syncthreads();
        if (k0 + BK < K) {
            fetch_lp<A_KM, TM>(A, lda, m0, k0 + BK, M, K, fa);
            fetch_lp<false, TN>(B, ldb, n0, k0 + BK, N, K, fb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int kq = kk + 8 * (lane >> 5);
            V ah[I], bh[J], al[I], bl[J];
#pragma unroll
            for (int i = 0; i < I; ++i) {
                const int r = wm * (TM / 2) + i * 32 + (lane & 31);
                ah[i] = *(const V*)&Ah[r][kq];
                if constexpr (X3) al[i] = *(const V*)&Al[r][kq];
            }
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int r = wn * (TN / 2) + j * 32 + (lane & 31);
                bh[j] = *(const V*)&Bh[r][kq];
                if constexpr (X3) bl[j] = *(const V*)&Bl[r][kq];
            }
#pragma unroll
            for (int i = 0; i < I; ++i)
#pragma unroll
                for (int j = 0; j < J; ++j) {
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
    for (int i = 0; i < I; ++i)
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int n = n0 + wn * (TN / 2) + j * 32 + (lane & 31);
            if (n >= N) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * (TM / 2) + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m < M) C[(int64_t)m * ldc + n] = acc[i][j][r];
            }
        }
}

// tile shape t: 0 = 128 x 128, 1 = 128 x 64, 2 = 64 x 128, 3 = 64 x 64 (TM x TN)
template <int TM, int TN>
hipError_t f32_launch(const float* A, int64_t lda, int a_km, const float* B, int64_t ldb, float* C, int64_t ldc,
                      int M, int N, int K, hipStream_t s) {
    const int nM = (M + TM - 1) / TM, nN = (N + TN - 1) / TN;
    dim3 grid(8 * ((nM * nN + 7) / 8));
    if (a_km)
        hipLaunchKernelGGL((gemm_f32_kernel<true, TM, TN>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K,
                           nM, nN);
    else
        hipLaunchKernelGGL((gemm_f32_kernel<false, TM, TN>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K,
                           nM, nN);
    return hipGetLastError();
}

hipError_t launch_gemm_f32(const float* A, int64_t lda, int a_km, const float* B, int64_t ldb,
                           float* C, int64_t ldc, int M, int N, int K, hipStream_t s, int tile) {
    if ((lda & 3) || (ldb & 3)) return hipErrorInvalidValue;
    switch (tile) {
        case 0: return f32_launch<128, 128>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case 1: return f32_launch<128, 64>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case 2: return f32_launch<64, 128>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case 3: return f32_launch<64, 64>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        default: return hipErrorInvalidValue;
    }
}

template <int PREC, int TM, int TN>
hipError_t lp_launch(const float* A, int64_t lda, int a_km, const float* B, int64_t ldb, float* C, int64_t ldc,
                     int M, int N, int K, hipStream_t s) {
    const int nM = (M + TM - 1) / TM, nN = (N + TN - 1) / TN;
    dim3 grid(8 * ((nM * nN + 7) / 8));
    if (a_km)
        hipLaunchKernelGGL((gemm_lp_kernel<PREC, true, TM, TN>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K,
                           nM, nN);
    else
        hipLaunchKernelGGL((gemm_lp_kernel<PREC, false, TM, TN>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N,
                           K, nM, nN);
    return hipGetLastError();
}

template <int PREC>
hipError_t lp_tile(const float* A, int64_t lda, int a_km, const float* B, int64_t ldb, float* C, int64_t ldc,
                   int M, int N, int K, hipStream_t s, int tile) {
    switch (tile) {
        case 0: return lp_launch<PREC, 128, 128>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case 1: return lp_launch<PREC, 128, 64>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case 2: return lp_launch<PREC, 64, 128>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        case 3: return lp_launch<PREC, 64, 64>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gemm_lp(int prec, const float* A, int64_t lda, int a_km, const float* B, int64_t ldb,
                          float* C, int64_t ldc, int M, int N, int K, hipStream_t s, int tile) {
    if ((lda & 3) || (ldb & 3)) return hipErrorInvalidValue;
    switch (prec) {
        case LSS_PREC_BF16: return lp_tile<LSS_PREC_BF16>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s, tile);
        case LSS_PREC_FP16: return lp_tile<LSS_PREC_FP16>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s, tile);
        case LSS_PREC_BF16X3: return lp_tile<LSS_PREC_BF16X3>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s, tile);
        case LSS_PREC_FP16X3: return lp_tile<LSS_PREC_FP16X3>(A, lda, a_km, B, ldb, C, ldc, M, N, K, s, tile);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lss
