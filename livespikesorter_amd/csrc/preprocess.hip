// Batch ingest, common-average reference and overlap-save FFT highpass for gfx950.
//
// Semantics (Kilosort4, kilosort/io.py:757-774): per channel the time mean of the padded window is
// removed, then per sample the lower median over channels (torch.median), then the circular FFT
// highpass X <- fftshift(ifft(fft(X) * conj(fft(f)))) with the 300 Hz filtfilt impulse response f
// padded/cropped to the window (kilosort/preprocessing.py:138-154).  That circular correlation is
// y[i] = sum_j g[j] x[(i+j) mod W] with g[j] = f[(W/2 + j) mod W]; g decays below 1.3e-13 (of a
// 0.98 centre tap) beyond |j| = 800, so the taps |j| <= 800 are applied with 8192-point FFTs in
// LDS (overlap-save), two channels per complex transform.
#include "lss_kernels.h"

namespace lss {

constexpr int kSumRows = 128;    // rows per channel-sum block

// ------------------------------------------------------------------------------------------------
// channel sums: int16 [W][C] -> int64 sums[C] (exact).  Each block of kSumRows rows writes its
// partial sums to part[block][C]; a second kernel adds the partials.  (One 64-bit atomic per
// channel and block instead serialised at the memory side: 40 us for a 60122-row window with
// 512-row blocks, and more with smaller blocks.)
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void channel_sums_kernel(const int16_t* __restrict__ win,
                                                            int64_t W, int C,
                                                            long long* __restrict__ part) {
    const int64_t r0 = (int64_t)blockIdx.x * kSumRows;
    const int64_t r1 = min(W, r0 + kSumRows);
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        long long acc = 0;
        for (int64_t r = r0; r < r1; ++r) acc += win[r * C + c];
        part[(int64_t)blockIdx.x * C + c] = acc;
    }
}

// C % 8 == 0: each thread owns one 16-byte column group (8 channels) and walks rows with stride
// G = 256 / (C/8); the G partial sums of a group are combined in LDS.
__global__ __launch_bounds__(256) void channel_sums_vec_kernel(const int16_t* __restrict__ win,
                                                               int64_t W, int C,
                                                               long long* __restrict__ part) {
    __shared__ int lpart[256][8];
    const int nv = C / 8;                 // vectors per row (<= 128)
    const int G = 256 / nv;               // row lanes
    const int tid = threadIdx.x;
    const int v = tid % nv, g = tid / nv;
    const int64_t r0 = (int64_t)blockIdx.x * kSumRows;
    const int64_t r1 = min(W, r0 + kSumRows);
    int acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (g < G) {
        const uint4* base = (const uint4*)win;
        int64_t r = r0 + g;
        for (; r + 3 * G < r1; r += 4 * G) {
            uint4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = base[(r + u * G) * nv + v];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t w4[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[2 * j] += (int16_t)(w4[j] & 0xffffu);
                    acc[2 * j + 1] += (int16_t)(w4[j] >> 16);
                }
            }
        }
        for (; r < r1; r += G) {
            const uint4 q = base[r * nv + v];
            const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[2 * j] += (int16_t)(w4[j] & 0xffffu);
                acc[2 * j + 1] += (int16_t)(w4[j] >> 16);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) lpart[tid][j] = acc[j];
    __syncthreads();
    if (tid < nv) {
        long long tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int gg = 0; gg < G; ++gg)
#pragma unroll
            for (int j = 0; j < 8; ++j) tot[j] += lpart[gg * nv + tid][j];
#pragma unroll
        for (int j = 0; j < 8; ++j) part[(int64_t)blockIdx.x * C + 8 * tid + j] = tot[j];
    }
}

// sums[c] = sum over blocks of part[b][c]: 16 channels x 16 block slices per workgroup
__global__ __launch_bounds__(256) void sums_reduce_kernel(const long long* __restrict__ part, int nb, int C,
                                                          long long* __restrict__ sums) {
    __shared__ long long red[16][17];
    const int tid = threadIdx.x, cl = tid & 15, sl = tid >> 4;
    const int c = blockIdx.x * 16 + cl;
    long long acc = 0;
    if (c < C)
#pragma unroll 4
        for (int b = sl; b < nb; b += 16) acc += part[(int64_t)b * C + c];
    red[sl][cl] = acc;
    __syncthreads();
    if (tid < 16 && c < C) {
        long long t = 0;
        for (int k = 0; k < 16; ++k) t += red[k][cl];
        sums[c] = t;
    }
}

int64_t channel_sum_blocks(int64_t W) { return (W + kSumRows - 1) / kSumRows; }

hipError_t launch_channel_sums(const int16_t* win, int64_t W, int C, long long* part, long long* sums,
                               hipStream_t s) {
    const int nb = (int)channel_sum_blocks(W);
    if (C % 8 == 0 && C / 8 <= 256 && (((uintptr_t)win) & 15) == 0)
        hipLaunchKernelGGL(channel_sums_vec_kernel, dim3(nb), dim3(256), 0, s, win, W, C, part);
    else
        hipLaunchKernelGGL(channel_sums_kernel, dim3(nb), dim3(256), 0, s, win, W, C, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sums_reduce_kernel, dim3((C + 15) / 16), dim3(256), 0, s, part, nb, C, sums);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// CAR: one wavefront per sample; exact lower median via binary search over order-preserving keys
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f2key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

// samples per block (8 waves x one pair): at 384 channels 26 KB of LDS, so more blocks share a CU
// than with 32 samples (86 vs 98 us at 384 x 60122; 64 samples: 174 us)
constexpr int kCarSamples = 16;

template <int VPL>
__global__ __launch_bounds__(512) void car_kernel(const int16_t* __restrict__ win, int64_t W, int C,
                                                  const long long* __restrict__ sums, int do_car,
                                                  float* __restrict__ out, int64_t ldo) {
    // tile[c][s] staged for a coalesced transposed store
    extern __shared__ float tile[];   // [C][kCarSamples + 1]
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t w0 = (int64_t)blockIdx.x * kCarSamples;
    const double invW = 1.0 / (double)W;
    float mean[VPL];
#pragma unroll
    for (int i = 0; i < VPL; ++i) {
        int c = lane + 64 * i;
        mean[i] = (c < C) ? (float)((double)sums[c] * invW) : 0.f;
    }
    const int k = (C - 1) / 2;   // lower median rank (torch.median)
    // two samples per step (si and si + 8), their loads and searches interleaved
    for (int si = wave; si < kCarSamples; si += 16) {
        const int64_t wa = w0 + si, wb = wa + 8;
        if (wa >= W) break;
        const bool hb = wb < W;
        float va[VPL], vb[VPL];
        uint32_t ka[VPL], kb[VPL];
#pragma unroll
        for (int i = 0; i < VPL; ++i) {
            const int c = lane + 64 * i;
            va[i] = (c < C) ? ((float)win[wa * C + c] - mean[i]) : 0.f;
            vb[i] = (c < C && hb) ? ((float)win[wb * C + c] - mean[i]) : 0.f;
            ka[i] = (c < C) ? f2key(va[i]) : 0xffffffffu;
            kb[i] = (c < C) ? f2key(vb[i]) : 0xffffffffu;
        }
        float meda = 0.f, medb = 0.f;
        if (do_car) {
            // the smallest key K with #{key <= K} >= k+1, one bit per step from the top: bit b
            // stays 0 when the largest key with the bits above fixed and bit b clear already
            // counts k+1
            uint32_t Ka = 0u, Kb = 0u;
#pragma unroll 2
            for (int b = 31; b >= 0; --b) {
                const uint32_t m = (1u << b) - 1u;
                const uint32_t ta = Ka | m, tb = Kb | m;
                int ca = 0, cb = 0;
#pragma unroll
                for (int i = 0; i < VPL; ++i) {
                    ca += __popcll(__ballot(ka[i] <= ta));
                    cb += __popcll(__ballot(kb[i] <= tb));
                }
                if (ca < k + 1) Ka |= 1u << b;
                if (cb < k + 1) Kb |= 1u << b;
            }
            meda = key2f(Ka);
            medb = key2f(Kb);
        }
#pragma unroll
        for (int i = 0; i < VPL; ++i) {
            const int c = lane + 64 * i;
            if (c < C) {
                tile[c * (kCarSamples + 1) + si] = va[i] - meda;
                if (hb) tile[c * (kCarSamples + 1) + si + 8] = vb[i] - medb;
            }
        }
    }
    __syncthreads();
    // store [C][kCarSamples] rows
    const int nvalid = (int)min<int64_t>(kCarSamples, W - w0);
    for (int idx = threadIdx.x; idx < C * kCarSamples; idx += blockDim.x) {
        int c = idx / kCarSamples, si = idx % kCarSamples;
        if (si < nvalid) out[(int64_t)c * ldo + w0 + si] = tile[c * (kCarSamples + 1) + si];
    }
}

template <int V>
hipError_t car_launch(dim3 grid, size_t lds, hipStream_t s, const int16_t* win, int64_t W, int C,
                      const long long* sums, int do_car, float* out, int64_t ldo) {
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)car_kernel<V>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(car_kernel<V>, grid, dim3(512), lds, s, win, W, C, sums, do_car, out, ldo);
    return hipGetLastError();
}

hipError_t launch_car_transpose(const int16_t* win, int64_t W, int C, const long long* sums,
                                int do_car, float* out, int64_t ldo, hipStream_t s) {
    if (C > 1024 || C < 1) return hipErrorInvalidValue;
    dim3 grid((unsigned)((W + kCarSamples - 1) / kCarSamples));
    size_t lds = sizeof(float) * (size_t)C * (kCarSamples + 1);
    int vpl = (C + 63) / 64;
    if (vpl <= 1) return car_launch<1>(grid, lds, s, win, W, C, sums, do_car, out, ldo);
    if (vpl <= 2) return car_launch<2>(grid, lds, s, win, W, C, sums, do_car, out, ldo);
    if (vpl <= 4) return car_launch<4>(grid, lds, s, win, W, C, sums, do_car, out, ldo);
    if (vpl <= 6) return car_launch<6>(grid, lds, s, win, W, C, sums, do_car, out, ldo);
    if (vpl <= 8) return car_launch<8>(grid, lds, s, win, W, C, sums, do_car, out, ldo);
    if (vpl <= 12) return car_launch<12>(grid, lds, s, win, W, C, sums, do_car, out, ldo);
    return car_launch<16>(grid, lds, s, win, W, C, sums, do_car, out, ldo);
}

// ------------------------------------------------------------------------------------------------
// 8192-point complex FFT in LDS: Stockham autosort, radix 8,8,8,8,2; 512 threads.
// ------------------------------------------------------------------------------------------------
constexpr int kFftThreads = 512;

// LDS index of FFT element i: 3 pad slots per 32 elements, so that the strided stores of the
// first radix-8 passes (stride 8 and 64 elements) spread over the banks (bank-conflict count of
// the five passes' loads and stores 40 -> 22, the minimum being 20).  71.7 KB per transform.
__device__ __forceinline__ int fpad(int i) { return i + 3 * (i >> 5); }
constexpr int kFftLds = 8192 + 3 * (8192 >> 5);

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i
__device__ __forceinline__ float2 cmi(float2 a) { return make_float2(a.y, -a.x); }

// 4-point DFT (forward, exp(-2 pi i nk/4)), outputs in natural order
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
    float2 a = cadd(x0, x2), b = cadd(x1, x3), c = csub(x0, x2), d = cmi(csub(x1, x3));
    x0 = cadd(a, b);
    x2 = csub(a, b);
    x1 = cadd(c, d);
    x3 = csub(c, d);
}

// 8-point DFT (forward), natural order in and out
__device__ __forceinline__ void dft8(float2* v) {
    const float r = 0.70710678118654752f;
    float2 u0 = cadd(v[0], v[4]), u1 = cadd(v[1], v[5]), u2 = cadd(v[2], v[6]), u3 = cadd(v[3], v[7]);
    float2 d0 = csub(v[0], v[4]), d1 = csub(v[1], v[5]), d2 = csub(v[2], v[6]), d3 = csub(v[3], v[7]);
    // d_n *= exp(-2 pi i n / 8)
    d1 = make_float2(r * (d1.x + d1.y), r * (d1.y - d1.x));
    d2 = cmi(d2);
    d3 = make_float2(r * (d3.y - d3.x), -r * (d3.x + d3.y));
    dft4(u0, u1, u2, u3);
    dft4(d0, d1, d2, d3);
    v[0] = u0; v[2] = u1; v[4] = u2; v[6] = u3;
    v[1] = d0; v[3] = d1; v[5] = d2; v[7] = d3;
}

// One Stockham pass of radix R over the LDS buffer (in place: read all, barrier, write all).
// Ns (the size of the sub-transforms already done) is a compile-time constant: the first pass
// (Ns = 1) has no twiddles, and a radix-8 pass loads only w, w^2 and w^4 from the table and forms
// w^3, w^5, w^6 and w^7 by complex products (at most two roundings from the table values).
template <int R, int Ns>
__device__ __forceinline__ void stockham_pass(float2* buf, const float2* __restrict__ tw) {
    constexpr int N = kFftN;
    constexpr int NB = N / R;                     // butterflies
    constexpr int PER = NB / kFftThreads;         // butterflies per thread
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = threadIdx.x + p * kFftThreads;
#pragma unroll
        for (int r = 0; r < R; ++r) v[p][r] = buf[fpad(j) + r * NB + 3 * ((r * NB) >> 5)];   // NB % 32 == 0
        if constexpr (Ns > 1) {
            const int step = (N / (Ns * R)) * (j % Ns);   // twiddle index multiplier
            if constexpr (R == 8) {
                const float2 w1 = tw[step], w2 = tw[2 * step], w4 = tw[4 * step];
                const float2 w3 = cmul(w1, w2);
                v[p][1] = cmul(v[p][1], w1);
                v[p][2] = cmul(v[p][2], w2);
                v[p][3] = cmul(v[p][3], w3);
                v[p][4] = cmul(v[p][4], w4);
                v[p][5] = cmul(v[p][5], cmul(w4, w1));
                v[p][6] = cmul(v[p][6], cmul(w4, w2));
                v[p][7] = cmul(v[p][7], cmul(w4, w3));
            } else {
                v[p][1] = cmul(v[p][1], tw[step]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = threadIdx.x + p * kFftThreads;
        const int k = j % Ns;
        if constexpr (R == 8) {
            dft8(v[p]);
        } else {
            float2 a = v[p][0], b = v[p][1];
            v[p][0] = cadd(a, b);
            v[p][1] = csub(a, b);
        }
        const int base = (j / Ns) * Ns * R + k;
        // fpad(base + r*Ns) with a constant offset per r: base is a multiple of 8 for Ns = 1, has
        // base % 32 < 8 for Ns = 8, and Ns is a multiple of 32 otherwise
        const int pb = fpad(base);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int off = (Ns == 1) ? r : (Ns == 8) ? 8 * r + 3 * (r >> 2) : r * Ns + 3 * ((r * Ns) >> 5);
            buf[pb + off] = v[p][r];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void fft8192(float2* buf, const float2* __restrict__ tw) {
    stockham_pass<8, 1>(buf, tw);
    stockham_pass<8, 8>(buf, tw);
    stockham_pass<8, 64>(buf, tw);
    stockham_pass<8, 512>(buf, tw);
    stockham_pass<2, 4096>(buf, tw);
}

__global__ __launch_bounds__(kFftThreads) void fft_rows_kernel(float2* data, const float2* __restrict__ tw) {
    extern __shared__ float2 fft_buf[];   // [kFftLds]
    float2* buf = fft_buf;
    float2* row = data + (int64_t)blockIdx.x * kFftN;
    for (int i = threadIdx.x; i < kFftN; i += kFftThreads) buf[fpad(i)] = row[i];
    __syncthreads();
    fft8192(buf, tw);
    for (int i = threadIdx.x; i < kFftN; i += kFftThreads) row[i] = buf[fpad(i)];
}

static hipError_t fft_lds_attr(const void* kernel) {
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(float2) * kFftLds));
}

hipError_t launch_fft_rows(float2* data, int nb, const float2* tw, hipStream_t s) {
    if (hipError_t e = fft_lds_attr((const void*)fft_rows_kernel)) return e;
    hipLaunchKernelGGL(fft_rows_kernel, dim3(nb), dim3(kFftThreads), sizeof(float2) * kFftLds, s, data, tw);
    return hipGetLastError();
}

// grid: (segments, channel pairs).  Segment b produces outputs [b*V, b*V+V) with V = N-Llo-Lhi.
__global__ __launch_bounds__(kFftThreads) void highpass_kernel(const float* __restrict__ x, int64_t ldx,
                                                               float* __restrict__ y, int64_t ldy,
                                                               int C, int64_t W, int Llo, int Lhi,
                                                               const float2* __restrict__ spec,
                                                               const float2* __restrict__ tw) {
    extern __shared__ float2 fft_buf[];   // [kFftLds]
    float2* buf = fft_buf;
    const int V = kFftN - Llo - Lhi;
    const int64_t o0 = (int64_t)blockIdx.x * V;
    const int ca = 2 * blockIdx.y;
    const int cb = ca + 1;
    const float* xa = x + (int64_t)ca * ldx;
    const float* xb = x + (int64_t)cb * ldx;
    const bool hasb = cb < C;
    // segment sample n <- window sample (o0 - Llo + n) mod W; the reduction is a subtraction
    // except where the segment wraps more than once (windows shorter than the FFT)
    int64_t start = o0 - Llo;
    start %= W;
    if (start < 0) start += W;
    const int Wi = (int)W, s0 = (int)start;
    for (int n = threadIdx.x; n < kFftN; n += kFftThreads) {
        int idx = s0 + n;
        if (idx >= Wi) {
            idx -= Wi;
            if (idx >= Wi) idx %= Wi;
        }
        buf[fpad(n)] = make_float2(xa[idx], hasb ? xb[idx] : 0.f);
    }
    __syncthreads();
    fft8192(buf, tw);
    // Z = conj(S * H); the forward FFT of Z conjugated is the (scaled) inverse transform
    for (int n = threadIdx.x; n < kFftN; n += kFftThreads) {
        float2 z = cmul(buf[fpad(n)], spec[n]);
        buf[fpad(n)] = make_float2(z.x, -z.y);
    }
    __syncthreads();
    fft8192(buf, tw);
    float* ya = y + (int64_t)ca * ldy;
    float* yb = y + (int64_t)cb * ldy;
    for (int i = threadIdx.x; i < V; i += kFftThreads) {
        int64_t o = o0 + i;
        if (o >= W) break;
        float2 r = buf[fpad(Llo + i)];
        ya[o] = r.x;
        if (hasb) yb[o] = -r.y;
    }
}

hipError_t launch_highpass(const float* x, int64_t ldx, float* y, int64_t ldy, int C, int64_t W,
                           int Llo, int Lhi, const float2* spec, const float2* tw, hipStream_t s) {
    int V = kFftN - Llo - Lhi;
    if (V <= 0) return hipErrorInvalidValue;
    int nseg = (int)((W + V - 1) / V);
    if (hipError_t e = fft_lds_attr((const void*)highpass_kernel)) return e;
    hipLaunchKernelGGL(highpass_kernel, dim3(nseg, (C + 1) / 2), dim3(kFftThreads), sizeof(float2) * kFftLds, s, x,
                       ldx, y, ldy, C, W, Llo, Lhi, spec, tw);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Spectrum of the overlap-save kernel for window length W, on the device in FP64.  The circular
// correlation kernel g[j] = f[(W/2 + j) mod W] (f = the taps padded / cropped to W as
// fft_highpass does, kilosort/preprocessing.py:138-154) enters the convolution as h[(-j) mod N],
// so spec[n] = (1/N) sum_{j=-Llo..Lhi} g[j] exp(+2 pi i n j / N), rounded to FP32 once.
// One thread per frequency; tw64[m] = exp(-2 pi i m / N) in FP64 (the exponent's index reduced
// mod N in integers, so every term uses an exactly rounded table value).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void hp_spectrum_kernel(const float* __restrict__ taps, int ft, int64_t W,
                                                          int Llo, int Lhi, const double2* __restrict__ tw64,
                                                          float2* __restrict__ spec) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= kFftN) return;
    double re = 0.0, im = 0.0;
    for (int j = -Llo; j <= Lhi; ++j) {
        const int64_t m = ((W / 2 + j) % W + W) % W;
        double g;
        if (ft < W) {
            const int64_t q = m - (W - ft) / 2;
            g = (q >= 0 && q < ft) ? (double)taps[q] : 0.0;
        } else if (ft > W) {
            g = (double)taps[(ft - W) / 2 + m];
        } else {
            g = (double)taps[m];
        }
        const int e = (int)(((int64_t)n * j % kFftN + kFftN) % kFftN);
        const double2 w = tw64[e];   // exp(-2 pi i e / N); the term needs its conjugate
        re += g * w.x;
        im -= g * w.y;
    }
    spec[n] = make_float2((float)(re / kFftN), (float)(im / kFftN));
}

hipError_t launch_hp_spectrum(const float* taps, int ft, int64_t W, int Llo, int Lhi, const double2* tw64,
                              float2* spec, hipStream_t s) {
    if (ft < 1 || W < 1 || Llo < 0 || Lhi < 0 || Llo + Lhi + 1 > kFftN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hp_spectrum_kernel, dim3(kFftN / 256), dim3(256), 0, s, taps, ft, W, Llo, Lhi, tw64, spec);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Window assembly from device samples: segments of int16 copied into window rows, all probes' in
// one launch.  Segment k writes n elements at dst from src (period 0: contiguous) or repeats the
// period-element row src (the rows before the recording starts).  Threads take 8-element chunks
// of the concatenated segments (16-byte accesses where both ends are 16-byte aligned).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gather_kernel(GatherArgs g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < g.chunks; q += stride) {
        int k = 0;
        while (k + 1 < g.n && g.seg[k + 1].chunk0 <= q) ++k;
        const GatherSeg& sg = g.seg[k];
        const int64_t i0 = (q - sg.chunk0) * 8;
        const int64_t m = sg.n - i0 < 8 ? sg.n - i0 : 8;
        if (sg.period == 0 && m == 8 && (((uintptr_t)(sg.dst + i0) | (uintptr_t)(sg.src + i0)) & 15) == 0) {
            *(int4*)(sg.dst + i0) = *(const int4*)(sg.src + i0);
        } else {
            for (int64_t i = i0; i < i0 + m; ++i) sg.dst[i] = sg.src[sg.period ? i % sg.period : i];
        }
    }
}

hipError_t launch_gather(const GatherSeg* segs, int n, hipStream_t s) {
    for (int k0 = 0; k0 < n; k0 += kGatherMaxSeg) {
        GatherArgs g{};
        g.n = std::min(kGatherMaxSeg, n - k0);
        int64_t c = 0;
        for (int k = 0; k < g.n; ++k) {
            g.seg[k] = segs[k0 + k];
            g.seg[k].chunk0 = c;
            c += (segs[k0 + k].n + 7) / 8;
        }
        g.chunks = c;
        if (c == 0) continue;
        const unsigned nb = (unsigned)std::min<int64_t>((c + 255) / 256, 2048);
        hipLaunchKernelGGL(gather_kernel, dim3(nb), dim3(256), 0, s, g);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace lss
