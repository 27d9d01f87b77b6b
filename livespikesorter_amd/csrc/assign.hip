// Feature-space cluster assignment of the reference's Python online sorter
// (src/Python/OnlineKilosort.py:211-284, cluster_centroids_pca_compute + assign_to_clusters).
//
//   score[i][k] = 2 <tF[i], mu[k]> - |mu[k]|^2          (tF[i], mu[k]: D = nC*K features)
//   label[i]    = argmax_k score[i][k] over clusters with (x_i - cx_k)^2 + (y_i - cy_k)^2 <= R^2,
//                 first maximum on ties, -1 when no cluster is in range (NaN centroids never are)
//
// A workgroup takes 64 spikes (their feature rows staged in LDS) and walks the clusters in tiles
// of 64 (centroid rows and norms staged in LDS); thread (s, q) scores spike s against the tile's
// clusters q, q + 4, ... and keeps its first maximum; the four partial winners of a spike are
// merged in cluster order at the end.  LDS rows are padded to an odd stride (conflict-free reads
// of a spike's row by consecutive lanes).  The distance test keeps the reference's rounding
// (dx^2 and dy^2 rounded, then added; no contraction into an FMA).
#include <math.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/lss.h"
#include "lss_kernels.h"

namespace lss {
lss_status set_error(lss_status st, const char* msg);

constexpr int kAsgS = 64;   // spikes per workgroup
constexpr int kAsgK = 64;   // clusters per LDS tile

__global__ __launch_bounds__(256) void assign_clusters_kernel(const float* __restrict__ tF, int64_t n, int D,
                                                              const float* __restrict__ mu, int K,
                                                              const float* __restrict__ xs,
                                                              const float* __restrict__ ys,
                                                              const float* __restrict__ cx,
                                                              const float* __restrict__ cy, float R2,
                                                              int* __restrict__ label) {
    extern __shared__ float asg_smem[];
    const int ld = D | 1;                      // odd row stride
    float* sX = asg_smem;                      // [kAsgS][ld]
    float* sM = sX + kAsgS * ld;               // [kAsgK][ld]
    float* sN = sM + kAsgK * ld;               // [kAsgK] |mu|^2
    float* sCx = sN + kAsgK;                   // [kAsgK]
    float* sCy = sCx + kAsgK;
    float* sBest = sCy + kAsgK;                // [4][kAsgS]
    int* sIdx = (int*)(sBest + 4 * kAsgS);     // [4][kAsgS]
    const int tid = threadIdx.x, s = tid & (kAsgS - 1), q = tid >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * kAsgS;
    for (int e = tid; e < kAsgS * D; e += 256) {
        const int r = e / D, d = e - r * D;
        sX[r * ld + d] = (i0 + r < n) ? tF[(i0 + r) * D + d] : 0.f;
    }
    const int64_t i = i0 + s;
    const float x = (i < n) ? xs[i] : 0.f, y = (i < n) ? ys[i] : 0.f;
    float best = -INFINITY;
    int bi = -1;
    for (int k0 = 0; k0 < K; k0 += kAsgK) {
        __syncthreads();                       // the previous tile is no longer read
        const int nk = min(kAsgK, K - k0);
        for (int e = tid; e < nk * D; e += 256) {
            const int r = e / D, d = e - r * D;
            sM[r * ld + d] = mu[(int64_t)(k0 + r) * D + d];
        }
        if (tid < nk) {
            sCx[tid] = cx[k0 + tid];
            sCy[tid] = cy[k0 + tid];
        }
        __syncthreads();
        if (tid < nk) {
            float acc = 0.f;   // squares rounded, then summed, as (mu**2).sum(1)
            for (int d = 0; d < D; ++d) acc = __fadd_rn(acc, __fmul_rn(sM[tid * ld + d], sM[tid * ld + d]));
            sN[tid] = acc;
        }
        __syncthreads();
        if (i < n) {
            for (int kk = q; kk < nk; kk += 4) {
                const float dx = __fsub_rn(x, sCx[kk]), dy = __fsub_rn(y, sCy[kk]);
                if (!(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)) <= R2)) continue;
                float dot = 0.f;
                for (int d = 0; d < D; ++d) dot = fmaf(sX[s * ld + d], sM[kk * ld + d], dot);
                const float sc = 2.f * dot - sN[kk];
                if (bi < 0 || sc > best) {     // ascending k: the first maximum stays
                    best = sc;
                    bi = k0 + kk;
                }
            }
        }
    }
    sBest[q * kAsgS + s] = best;
    sIdx[q * kAsgS + s] = bi;
    __syncthreads();
    if (q == 0 && i < n) {
        float b = -INFINITY;
        int l = -1;
        for (int p = 0; p < 4; ++p) {
            const int j = sIdx[p * kAsgS + s];
            if (j < 0) continue;
            const float v = sBest[p * kAsgS + s];
            if (l < 0 || v > b || (v == b && j < l)) {
                b = v;
                l = j;
            }
        }
        label[i] = l;
    }
}

hipError_t launch_assign_clusters(const float* tF, int64_t n, int D, const float* mu, int K, const float* xs,
                                  const float* ys, const float* cx, const float* cy, float R2, int* label,
                                  hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (D < 1 || D > 256 || K < 1) return hipErrorInvalidValue;
    const int ld = D | 1;
    const size_t lds =
        sizeof(float) * ((size_t)(kAsgS + kAsgK) * ld + 3 * kAsgK + 4 * kAsgS) + sizeof(int) * 4 * kAsgS;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)assign_clusters_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const unsigned grid = (unsigned)((n + kAsgS - 1) / kAsgS);
    hipLaunchKernelGGL(assign_clusters_kernel, dim3(grid), dim3(256), lds, s, tF, n, D, mu, K, xs, ys, cx, cy, R2,
                       label);
    return hipGetLastError();
}

// centroid[t][j*K + k] = Ucc[j][t][k]: each template's PC features on its nC nearest channels
// (cluster_centroids_pca_compute with pc_feature_ind[t] = iCC[:, iU[t]])
__global__ void centroids_kernel(const float* __restrict__ Ucc, int T, int nC, int K, float* __restrict__ out) {
    const int64_t total = (int64_t)T * nC * K;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t t = e / (nC * K);
        const int r = (int)(e - t * nC * K), j = r / K, k = r - j * K;
        out[e] = Ucc[((int64_t)j * T + t) * K + k];
    }
}

hipError_t launch_cluster_centroids(const float* Ucc, int T, int nC, int K, float* out, hipStream_t s) {
    const int64_t total = (int64_t)T * nC * K;
    const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 1024);
    hipLaunchKernelGGL(centroids_kernel, dim3(grid), dim3(256), 0, s, Ucc, T, nC, K, out);
    return hipGetLastError();
}

}  // namespace lss

extern "C" lss_status lss_assign_clusters(const float* d_tF, int64_t n, int32_t D, const float* d_mu,
                                          int32_t n_clusters, const float* d_x, const float* d_y, const float* d_cx,
                                          const float* d_cy, double R, int32_t* d_label, void* stream) {
    if (n < 0 || (n > 0 && (!d_tF || !d_x || !d_y || !d_label)) || D < 1 || D > 256 || n_clusters < 1 || !d_mu ||
        !d_cx || !d_cy || !std::isfinite(R))
        return lss::set_error(LSS_ERR_ARG, "lss_assign_clusters: bad argument");
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = lss::launch_assign_clusters(d_tF, n, D, d_mu, n_clusters, d_x, d_y, d_cx, d_cy, (float)(R * R),
                                               d_label, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return lss::set_error(LSS_ERR_HIP, (std::string("lss_assign_clusters: ") + hipGetErrorString(e)).c_str());
    return LSS_OK;
}
