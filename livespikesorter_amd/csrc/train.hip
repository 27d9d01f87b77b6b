// Training-side artifacts of Kilosort4, built in FP64 on the GPU:
//   FP64 GEMM with split K (covariance of filtered batches, drift products, Newton-Schulz inverse)
//   batched parallel-order Jacobi eigensolver in LDS (local ZCA whitening rows, full ZCA, PCA)
//   Gaussian interpolation kernel of the drift model
//   isolated single-channel peaks, ordered snippet compaction and snippet normalisation (wPCA)
// The Python the artifacts follow: kilosort/preprocessing.py:8-69, 96-113 (whitening, drift
// matrix), kilosort/datashift.py:176-184, 205-219 (kernel2D, iKxx), kilosort/spikedetect.py:17-79
// (isolated peaks, snippets, TruncatedSVD wPCA).
#include "lss_kernels.h"

namespace lss {

// ---- FP64 GEMM ----------------------------------------------------------------------------------
// P[z][m][n] = sum_{k in chunk z} A(m,k) B(k,n) with A(m,k) = A[m*lda + k] and
// B(k,n) = B[n*ldb + k] (BNK) or B[k*ldb + n].  64x64 output tiles, 32-deep K steps staged in LDS
// as FP64 (FP32 inputs are widened exactly), 4x4 outputs per thread strided by 16 so that a
// wavefront's LDS reads are 4 broadcast A values and 16 consecutive B values.  sym: the output is
// symmetric (A = B); tiles below the diagonal are skipped and read mirrored by the reduction.
constexpr int kG = 64, kGK = 32;

template <typename TA, typename TB, bool BNK>
__global__ __launch_bounds__(256) void gemm64_kernel(const TA* __restrict__ A, int64_t lda, const TB* __restrict__ B,
                                                     int64_t ldb, double* __restrict__ P, int M, int N, int64_t K,
                                                     int64_t kchunk, int sym) {
    const int tm = blockIdx.y, tn = blockIdx.x;
    if (sym && tn < tm) return;
    __shared__ double As[kGK][kG + 1];
    __shared__ double Bs[kGK][kG + 1];
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const int64_t k0 = (int64_t)blockIdx.z * kchunk;
    const int64_t k1 = k0 + kchunk < K ? k0 + kchunk : K;
    const int m0 = tm * kG, n0 = tn * kG;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int64_t kb = k0; kb < k1; kb += kGK) {
        {
            const int r = t >> 2, kk = (t & 3) * 8;
            const int m = m0 + r;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t k = kb + kk + e;
                As[kk + e][r] = (m < M && k < k1) ? (double)A[(int64_t)m * lda + k] : 0.0;
            }
        }
        if (BNK) {
            const int r = t >> 2, kk = (t & 3) * 8;
            const int n = n0 + r;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t k = kb + kk + e;
                Bs[kk + e][r] = (n < N && k < k1) ? (double)B[(int64_t)n * ldb + k] : 0.0;
            }
        } else {
            const int kk = t >> 3, nn = (t & 7) * 8;
            const int64_t k = kb + kk;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int n = n0 + nn + e;
                Bs[kk][nn + e] = (n < N && k < k1) ? (double)B[k * ldb + n] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < kGK; ++k) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[k][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    double* p = P + (int64_t)blockIdx.z * M * N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = m0 + ty + 16 * i, n = n0 + tx + 16 * j;
            if (m < M && n < N) p[(int64_t)m * N + n] = acc[i][j];
        }
}

// out[m][n] = beta*out[m][n] + alpha * sum_z P[z][m][n] (+ diag when m == n), chunks summed in
// order z = 0, 1, ...; with sym the entries of skipped tiles are read from the mirrored tile.
__global__ __launch_bounds__(256) void reduce64_kernel(const double* __restrict__ P, int S, int M, int N, int sym,
                                                       double alpha, double beta, double diag, double* __restrict__ out,
                                                       int64_t ldo) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)M * N) return;
    const int m = (int)(idx / N), n = (int)(idx % N);
    int mm = m, nn = n;
    if (sym && (n / kG) < (m / kG)) { mm = n; nn = m; }
    double acc = 0.0;
    for (int z = 0; z < S; ++z) acc += P[((int64_t)z * M + mm) * N + nn];
    double v = alpha * acc;
    if (beta != 0.0) v += beta * out[(int64_t)m * ldo + n];
    if (m == n) v += diag;
    out[(int64_t)m * ldo + n] = v;
}

// Split count that gives about 1024 workgroups, and the chunk length (a multiple of kGK).
void gemm64_split(int M, int N, int64_t K, int sym, int* S, int64_t* kchunk) {
    const int64_t tm = (M + kG - 1) / kG, tn = (N + kG - 1) / kG;
    const int64_t tiles = sym ? tm * (tm + 1) / 2 : tm * tn;
    int64_t s = (1024 + tiles - 1) / tiles;
    const int64_t kmax = (K + kGK - 1) / kGK;
    if (s > kmax) s = kmax;
    if (s < 1) s = 1;
    int64_t ch = (K + s - 1) / s;
    ch = (ch + kGK - 1) / kGK * kGK;
    if (ch < kGK) ch = kGK;
    *S = (int)((K + ch - 1) / ch);
    if (*S < 1) *S = 1;
    *kchunk = ch;
}

template <typename TA, typename TB, bool BNK>
static hipError_t gemm64_launch(const TA* A, int64_t lda, const TB* B, int64_t ldb, double* P, int M, int N, int64_t K,
                                int S, int64_t kchunk, int sym, hipStream_t s) {
    const dim3 grid((N + kG - 1) / kG, (M + kG - 1) / kG, S);
    hipLaunchKernelGGL((gemm64_kernel<TA, TB, BNK>), grid, dim3(256), 0, s, A, lda, B, ldb, P, M, N, K, kchunk, sym);
    return hipGetLastError();
}

hipError_t launch_gemm64_f32_nk(const float* A, int64_t lda, const float* B, int64_t ldb, double* P, int M, int N,
                                int64_t K, int S, int64_t kchunk, int sym, hipStream_t s) {
    return gemm64_launch<float, float, true>(A, lda, B, ldb, P, M, N, K, S, kchunk, sym, s);
}
hipError_t launch_gemm64_nk(const double* A, int64_t lda, const double* B, int64_t ldb, double* P, int M, int N,
                            int64_t K, int S, int64_t kchunk, int sym, hipStream_t s) {
    return gemm64_launch<double, double, true>(A, lda, B, ldb, P, M, N, K, S, kchunk, sym, s);
}
hipError_t launch_gemm64_kn(const double* A, int64_t lda, const double* B, int64_t ldb, double* P, int M, int N,
                            int64_t K, int S, int64_t kchunk, hipStream_t s) {
    return gemm64_launch<double, double, false>(A, lda, B, ldb, P, M, N, K, S, kchunk, 0, s);
}

hipError_t launch_reduce64(const double* P, int S, int M, int N, int sym, double alpha, double beta, double diag,
                           double* out, int64_t ldo, hipStream_t s) {
    const int64_t n = (int64_t)M * N;
    hipLaunchKernelGGL(reduce64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, S, M, N, sym, alpha, beta,
                       diag, out, ldo);
    return hipGetLastError();
}

// *out = sum of squares of the [M][N] matrix x (row stride ldx), one workgroup, fixed order.
__global__ __launch_bounds__(1024) void sumsq64_kernel(const double* __restrict__ x, int M, int N, int64_t ldx,
                                                       double* __restrict__ out) {
    __shared__ double red[1024];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < (int64_t)M * N; i += 1024) {
        const double v = x[(i / N) * ldx + i % N];
        acc = fma(v, v, acc);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

hipError_t launch_sumsq64(const double* x, int M, int N, int64_t ldx, double* out, hipStream_t s) {
    hipLaunchKernelGGL(sumsq64_kernel, dim3(1), dim3(1024), 0, s, x, M, N, ldx, out);
    return hipGetLastError();
}

// x[i][j] = (double)src[i][j] conversions and the reverse, for the whitening operator
__global__ __launch_bounds__(256) void f64_to_f32_kernel(const double* __restrict__ x, int M, int N, int64_t ldx,
                                                         float* __restrict__ y, int64_t ldy) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)M * N) return;
    const int64_t i = idx / N, j = idx % N;
    y[i * ldy + j] = (float)x[i * ldx + j];
}

hipError_t launch_f64_to_f32(const double* x, int M, int N, int64_t ldx, float* y, int64_t ldy, hipStream_t s) {
    const int64_t n = (int64_t)M * N;
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, M, N, ldx, y, ldy);
    return hipGetLastError();
}

// ---- drift interpolation kernel -----------------------------------------------------------------
// K[i][j] = exp(-((xc_i - xc_j)^2 + (yc_i - shift_i - yc_j)^2) / (2 sig^2)) + lam [i == j]
// (kernel2D of kilosort/datashift.py:176-184 between the shifted sites and the sites).
__global__ __launch_bounds__(256) void interp_kernel_kernel(const double* __restrict__ xc, const double* __restrict__ yc,
                                                            const double* __restrict__ shift, int C, double sig,
                                                            double lam, double* __restrict__ Kout) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)C * C) return;
    const int i = (int)(idx / C), j = (int)(idx % C);
    const double yi = shift ? yc[i] - shift[i] : yc[i];
    const double dx = xc[i] - xc[j], dy = yi - yc[j];
    const double ds = dx * dx + dy * dy;
    double v = exp(-ds / (2.0 * sig * sig));
    if (i == j) v += lam;
    Kout[idx] = v;
}

hipError_t launch_interp_kernel(const double* xc, const double* yc, const double* shift, int C, double sig, double lam,
                                double* Kout, hipStream_t s) {
    const int64_t n = (int64_t)C * C;
    hipLaunchKernelGGL(interp_kernel_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xc, yc, shift, C, sig,
                       lam, Kout);
    return hipGetLastError();
}

// ---- batched Jacobi eigensolver -----------------------------------------------------------------
// One workgroup per symmetric n x n matrix (n <= kEigMax), padded to even np with a zero row and
// column.  Parallel (round-robin tournament) ordering: every step rotates np/2 disjoint pairs; a
// sweep of np-1 steps visits every pair once.  Rotation of pair (p, q) with the classical angle
// (tau = (a_qq - a_pp) / (2 a_pq), t = sign(tau) / (|tau| + sqrt(1 + tau^2))), applied to the rows
// then the columns of A and the columns of V; the rotated off-diagonal a_pq is set to zero.
// Sweeps stop when the off-diagonal sum of squares is <= 1e-30 of the total (or after kEigSweeps).
constexpr int kEigSweeps = 40;

__device__ inline void rr_pair(int i, int step, int np, int& a, int& b) {
    const int n1 = np - 1;
    if (i == 0) {
        a = 0;
        b = (step % n1) + 1;
    } else {
        a = ((i + step) % n1) + 1;                 // position 1 + i
        b = ((np - i - 1 + step) % n1) + 1;        // position np - i
    }
}

struct EigLds {
    double* A;     // [np][ld]
    double* V;     // [np][ld]
    double* rc;    // [np/2]
    double* rs;
    int* rp;
    int* rq;
    double* red;   // [512]
    int np, ld;
};

__device__ EigLds eig_lds(double* base, int np) {
    EigLds L;
    L.np = np;
    L.ld = np + 1;
    L.A = base;
    L.V = L.A + (size_t)np * L.ld;
    L.red = L.V + (size_t)np * L.ld;
    L.rc = L.red + 512;
    L.rs = L.rc + np / 2;
    L.rp = (int*)(L.rs + np / 2);
    L.rq = L.rp + np / 2;
    return L;
}

size_t eig_lds_bytes(int np) {
    const int ld = np + 1;
    return sizeof(double) * (2 * (size_t)np * ld + 512 + np) + sizeof(int) * np;
}

__device__ void jacobi_eig(EigLds& L) {
    const int t = threadIdx.x, nth = blockDim.x;
    const int np = L.np, ld = L.ld, m = np / 2;
    double* A = L.A;
    double* V = L.V;
    for (int e = t; e < np * np; e += nth) V[(e / np) * ld + e % np] = (e / np == e % np) ? 1.0 : 0.0;
    __syncthreads();
    for (int sweep = 0; sweep < kEigSweeps; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int e = t; e < np * np; e += nth) {
            const int i = e / np, j = e % np;
            const double v = A[i * ld + j];
            tot = fma(v, v, tot);
            if (i != j) off = fma(v, v, off);
        }
        L.red[t] = off;
        L.red[256 + t] = tot;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (t < st) {
                L.red[t] += L.red[t + st];
                L.red[256 + t] += L.red[256 + t + st];
            }
            __syncthreads();
        }
        const bool done = L.red[0] <= 1e-30 * L.red[256];
        __syncthreads();
        if (done) break;
        for (int step = 0; step < np - 1; ++step) {
            if (t < m) {
                int a, b;
                rr_pair(t, step, np, a, b);
                const int p = a < b ? a : b, q = a < b ? b : a;
                const double apq = A[p * ld + q];
                double c = 1.0, s = 0.0;
                if (apq != 0.0) {
                    const double tau = (A[q * ld + q] - A[p * ld + p]) / (2.0 * apq);
                    const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + tt * tt);
                    s = tt * c;
                }
                L.rc[t] = c;
                L.rs[t] = s;
                L.rp[t] = p;
                L.rq[t] = q;
            }
            __syncthreads();
            for (int e = t; e < m * np; e += nth) {   // rows: A <- J^T A
                const int i = e / np, j = e % np;
                const int p = L.rp[i], q = L.rq[i];
                const double c = L.rc[i], s = L.rs[i];
                const double x = A[p * ld + j], y = A[q * ld + j];
                A[p * ld + j] = c * x - s * y;
                A[q * ld + j] = s * x + c * y;
            }
            __syncthreads();
            for (int e = t; e < m * np; e += nth) {   // columns: A <- A J, V <- V J
                const int i = e % m, j = e / m;
                const int p = L.rp[i], q = L.rq[i];
                const double c = L.rc[i], s = L.rs[i];
                const double x = A[j * ld + p], y = A[j * ld + q];
                A[j * ld + p] = (j == q) ? 0.0 : c * x - s * y;
                A[j * ld + q] = (j == p) ? 0.0 : s * x + c * y;
                const double vx = V[j * ld + p], vy = V[j * ld + q];
                V[j * ld + p] = c * vx - s * vy;
                V[j * ld + q] = s * vx + c * vy;
            }
            __syncthreads();
        }
    }
}

// Local ZCA rows (kilosort/preprocessing.py:19-35): workgroup j takes CC over its neighbour list
// idx[j][0..nr), and writes row j of W: W[j][idx[j][i]] = (E diag(1/sqrt(|d| + eps)) E^T)[0][i],
// zero elsewhere (the SVD of the symmetric PSD local CC gives |eigenvalues| as singular values).
__global__ __launch_bounds__(256) void whiten_local_kernel(const double* __restrict__ cc, int64_t ldcc,
                                                           const int* __restrict__ idx, int nr, int C, double eps,
                                                           double* __restrict__ W, int64_t ldw) {
    extern __shared__ double lds[];
    const int np = nr + (nr & 1);
    EigLds L = eig_lds(lds, np);
    double* row = (double*)(L.rq + np / 2);       // [C]
    const int j = blockIdx.x, t = threadIdx.x;
    const int* ix = idx + (int64_t)j * nr;
    for (int e = t; e < np * np; e += 256) {
        const int a = e / np, b = e % np;
        L.A[a * L.ld + b] = (a < nr && b < nr) ? cc[(int64_t)ix[a] * ldcc + ix[b]] : 0.0;
    }
    for (int c = t; c < C; c += 256) row[c] = 0.0;
    __syncthreads();
    jacobi_eig(L);
    if (t < nr) {
        double acc = 0.0;
        for (int k = 0; k < np; ++k) {
            const double w = 1.0 / sqrt(fabs(L.A[k * L.ld + k]) + eps);
            acc = fma(L.V[0 * L.ld + k] * L.V[t * L.ld + k], w, acc);
        }
        row[ix[t]] = acc;
    }
    __syncthreads();
    for (int c = t; c < C; c += 256) W[(int64_t)j * ldw + c] = row[c];
}

// Full ZCA of one n x n matrix (kilosort/preprocessing.py:8-17).
__global__ __launch_bounds__(256) void whiten_full_kernel(const double* __restrict__ cc, int64_t ldcc, int n, double eps,
                                                          double* __restrict__ W, int64_t ldw) {
    extern __shared__ double lds[];
    const int np = n + (n & 1);
    EigLds L = eig_lds(lds, np);
    const int t = threadIdx.x;
    for (int e = t; e < np * np; e += 256) {
        const int a = e / np, b = e % np;
        L.A[a * L.ld + b] = (a < n && b < n) ? cc[(int64_t)a * ldcc + b] : 0.0;
    }
    __syncthreads();
    jacobi_eig(L);
    for (int e = t; e < n * n; e += 256) {
        const int a = e / n, b = e % n;
        double acc = 0.0;
        for (int k = 0; k < np; ++k) {
            const double w = 1.0 / sqrt(fabs(L.A[k * L.ld + k]) + eps);
            acc = fma(L.V[a * L.ld + k] * L.V[b * L.ld + k], w, acc);
        }
        W[(int64_t)a * ldw + b] = acc;
    }
}

// Leading eigenvectors of a symmetric PSD n x n Gram matrix (right singular vectors of the data):
// out[r][m] = +-V[m][k_r] for the npc largest eigenvalues (ties: lower index first), each signed so
// that its largest-magnitude entry (first on ties) is positive (sklearn's svd_flip with
// u_based_decision=False, which TruncatedSVD applies); sv[r] = sqrt(max(d, 0)).
__global__ __launch_bounds__(256) void pca_kernel(const double* __restrict__ G, int64_t ldg, int n, int npc,
                                                  double* __restrict__ out, double* __restrict__ sv) {
    extern __shared__ double lds[];
    const int np = n + (n & 1);
    EigLds L = eig_lds(lds, np);
    int* ord = (int*)(L.rq + np / 2);             // [npc]
    double* sgn = (double*)(ord + 128);           // [npc]
    const int t = threadIdx.x;
    for (int e = t; e < np * np; e += 256) {
        const int a = e / np, b = e % np;
        L.A[a * L.ld + b] = (a < n && b < n) ? G[(int64_t)a * ldg + b] : 0.0;
    }
    __syncthreads();
    jacobi_eig(L);
    if (t == 0) {
        // the zero padding column is an exact eigenvector; it is never selected ahead of a real one
        for (int r = 0; r < npc; ++r) {
            int best = -1;
            for (int k = 0; k < n; ++k) {
                bool used = false;
                for (int u = 0; u < r; ++u) used |= (ord[u] == k);
                if (used) continue;
                if (best < 0 || L.A[k * L.ld + k] > L.A[best * L.ld + best]) best = k;
            }
            ord[r] = best;
            int am = 0;
            double vm = -1.0;
            for (int mm = 0; mm < n; ++mm) {
                const double v = fabs(L.V[mm * L.ld + best]);
                if (v > vm) { vm = v; am = mm; }
            }
            sgn[r] = L.V[am * L.ld + best] < 0.0 ? -1.0 : 1.0;
        }
    }
    __syncthreads();
    for (int e = t; e < npc * n; e += 256) {
        const int r = e / n, mm = e % n;
        out[(int64_t)r * n + mm] = sgn[r] * L.V[mm * L.ld + ord[r]];
    }
    if (t < npc) {
        const double d = L.A[ord[t] * L.ld + ord[t]];
        sv[t] = sqrt(d > 0.0 ? d : 0.0);
    }
}

static hipError_t set_lds(const void* fn, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_whiten_local(const double* cc, int64_t ldcc, const int* idx, int nr, int C, double eps, double* W,
                               int64_t ldw, hipStream_t s) {
    if (nr < 1 || nr > kEigMax || C < 1 || C > 1024) return hipErrorInvalidValue;
    const int np = nr + (nr & 1);
    const size_t bytes = eig_lds_bytes(np) + sizeof(double) * C;
    hipError_t e = set_lds((const void*)whiten_local_kernel, bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(whiten_local_kernel, dim3(C), dim3(256), bytes, s, cc, ldcc, idx, nr, C, eps, W, ldw);
    return hipGetLastError();
}

hipError_t launch_whiten_full(const double* cc, int64_t ldcc, int n, double eps, double* W, int64_t ldw, hipStream_t s) {
    if (n < 1 || n > kEigMax) return hipErrorInvalidValue;
    const int np = n + (n & 1);
    const size_t bytes = eig_lds_bytes(np);
    hipError_t e = set_lds((const void*)whiten_full_kernel, bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(whiten_full_kernel, dim3(1), dim3(256), bytes, s, cc, ldcc, n, eps, W, ldw);
    return hipGetLastError();
}

hipError_t launch_pca(const double* G, int64_t ldg, int n, int npc, double* out, double* sv, hipStream_t s) {
    if (n < 1 || n > kEigMax || npc < 1 || npc > n || npc > 128) return hipErrorInvalidValue;
    const int np = n + (n & 1);
    const size_t bytes = eig_lds_bytes(np) + sizeof(int) * 128 + sizeof(double) * 128;
    hipError_t e = set_lds((const void*)pca_kernel, bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pca_kernel, dim3(1), dim3(256), bytes, s, G, ldg, n, npc, out, sv);
    return hipGetLastError();
}

// ---- isolated single-channel peaks (kilosort/spikedetect.py:17-47) ------------------------------
// ispeak(c,t) = |X(c,t)| equals the maximum of |X| over channels [c-lc, c+lc] x times [t-lt, t+lt]
// (max_pool2d, out-of-range cells excluded) and |X(c,t)| > th.  X is the [W][ldx] time-major
// window; the flags go to a channel-major bitmap pk[c][word], bit b of word w = time 32w+b.
constexpr int kPkT = 256;      // times per workgroup (one per thread)
constexpr int kPkC = 8;        // channels per workgroup
constexpr int kPkHalo = 8;     // max supported lc and lt

__global__ __launch_bounds__(kPkT) void ispeak_kernel(const float* __restrict__ X, int64_t ldx, int C, int64_t W, int lc,
                                                      int lt, float th, unsigned* __restrict__ pk, int64_t nw) {
    __shared__ float ax[kPkT + 2 * kPkHalo][kPkC + 2 * kPkHalo + 1];
    const int t = threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.x * kPkT;
    const int c0 = blockIdx.y * kPkC;
    const int rows = kPkT + 2 * lt, cols = kPkC + 2 * lc;
    for (int e = t; e < rows * cols; e += kPkT) {
        const int r = e / cols, cc = e % cols;
        const int64_t tt = t0 - lt + r;
        const int ch = c0 - lc + cc;
        ax[r][cc] = (tt >= 0 && tt < W && ch >= 0 && ch < C) ? fabsf(X[tt * ldx + ch]) : -1.0f;
    }
    __syncthreads();
    // time maxima per staged channel, then channel maxima; the unrolled loops keep tmax in registers
    float tmax[kPkC + 2 * kPkHalo];
#pragma unroll
    for (int cc = 0; cc < kPkC + 2 * kPkHalo; ++cc) {
        float v = -1.0f;
        if (cc < cols)
            for (int r = 0; r <= 2 * lt; ++r) v = fmaxf(v, ax[t + r][cc]);
        tmax[cc] = v;
    }
    const int64_t tt = t0 + t;
#pragma unroll
    for (int k = 0; k < kPkC; ++k) {
        const int ch = c0 + k;
        float v = -1.0f;
#pragma unroll
        for (int d = 0; d <= 2 * kPkHalo; ++d)
            if (d <= 2 * lc) v = fmaxf(v, tmax[k + d]);
        const float a = ax[t + lt][k + lc];
        const bool pkf = (ch < C) && (tt < W) && (v == a) && (a > th);
        const unsigned long long bits = __ballot(pkf);
        // wavefront lanes cover 64 consecutive times: two words of channel ch
        const int lane = t & 63;
        if (ch < C && lane < 2) {
            const int64_t word = (t0 + (t & ~63)) / 32 + lane;
            if (word < nw) pk[(int64_t)ch * nw + word] = (unsigned)(bits >> (32 * lane));
        }
    }
}

// Bits [lo, hi] (clipped to [0, 32*nw)) of one bitmap row, counted.
__device__ inline int count_bits(const unsigned* __restrict__ row, int64_t nw, int64_t lo, int64_t hi) {
    if (lo < 0) lo = 0;
    if (hi > 32 * nw - 1) hi = 32 * nw - 1;
    if (hi < lo) return 0;
    int n = 0;
    for (int64_t w = lo >> 5; w <= (hi >> 5); ++w) {
        unsigned m = row[w];
        const int64_t b0 = w * 32;
        if (lo > b0) m &= ~0u << (lo - b0);
        if (hi < b0 + 31) m &= ~0u >> (b0 + 31 - hi);
        n += __popc(m);
    }
    return n;
}

// iso(c,t) = ispeak(c,t) and exactly one flag in channels [c-gc, c+gc] x times [t-gt, t+gt]
// (the avg_pool2d count of spikedetect.py:24-30, 36-37), cleared for t < edge or t >= W - edge.
__global__ __launch_bounds__(256) void isolated_kernel(const unsigned* __restrict__ pk, int C, int64_t W, int64_t nw,
                                                       int gc, int gt, int edge, unsigned* __restrict__ iso) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)C * nw) return;
    const int c = (int)(idx / nw);
    const int64_t w = idx % nw;
    unsigned bits = pk[idx], out = 0u;
    while (bits) {
        const int b = __ffs(bits) - 1;
        bits &= bits - 1;
        const int64_t tt = w * 32 + b;
        if (tt < edge || tt >= W - edge) continue;
        int n = 0;
        for (int cc = c - gc; cc <= c + gc; ++cc) {
            if (cc < 0 || cc >= C) continue;
            n += count_bits(pk + (int64_t)cc * nw, nw, tt - gt, tt + gt);
        }
        if (n == 1) out |= 1u << b;
    }
    iso[idx] = out;
}

// Ordered compaction of the iso bitmap in (channel, time) order (torch's nonzero order):
// block sums, one-workgroup scan of the sums, then per-block rescan and snippet writes.
constexpr int kScanB = 1024;

__global__ __launch_bounds__(kScanB) void popcount_blocks_kernel(const unsigned* __restrict__ iso, int64_t nwords,
                                                                 int* __restrict__ bsum) {
    __shared__ int red[kScanB];
    const int64_t i = (int64_t)blockIdx.x * kScanB + threadIdx.x;
    red[threadIdx.x] = i < nwords ? __popc(iso[i]) : 0;
    __syncthreads();
    for (int st = kScanB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// Exclusive scan of nb block sums in place (one workgroup, serial chunks); total into *total.
__global__ __launch_bounds__(kScanB) void scan_blocks_kernel(int* __restrict__ bsum, int nb, long long* __restrict__ total) {
    __shared__ long long sh[kScanB];
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += kScanB) {
        const int i = base + threadIdx.x;
        const long long v = i < nb ? bsum[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kScanB; off <<= 1) {
            const long long add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nb) bsum[i] = (int)(carry + sh[threadIdx.x] - v);
        __syncthreads();
        if (threadIdx.x == kScanB - 1) carry += sh[kScanB - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanB) void emit_snippets_kernel(const unsigned* __restrict__ iso, int64_t nwords,
                                                               int64_t nw, const int* __restrict__ boff,
                                                               const float* __restrict__ X, int64_t ldx, int nt,
                                                               int twav_min, float* __restrict__ clips,
                                                               int64_t max_clips, int* __restrict__ chan,
                                                               int* __restrict__ tim) {
    __shared__ int sh[kScanB];
    const int64_t i = (int64_t)blockIdx.x * kScanB + threadIdx.x;
    const unsigned bits0 = i < nwords ? iso[i] : 0u;
    const int v = __popc(bits0);
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanB; off <<= 1) {
        const int add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t pos = (int64_t)boff[blockIdx.x] + sh[threadIdx.x] - v;
    if (!bits0) return;
    const int c = (int)(i / nw);
    const int64_t w = i % nw;
    unsigned bits = bits0;
    while (bits) {
        const int b = __ffs(bits) - 1;
        bits &= bits - 1;
        if (pos < max_clips) {
            const int64_t tt = w * 32 + b;
            float* dst = clips + pos * nt;
            for (int m = 0; m < nt; ++m) dst[m] = X[(tt - twav_min + m) * ldx + c];
            if (chan) chan[pos] = c;
            if (tim) tim[pos] = (int)tt;
        }
        ++pos;
    }
}

hipError_t launch_isolated_peaks(const float* X, int64_t ldx, int C, int64_t W, int lc, int lt, float th, int gc,
                                 int gt, int edge, unsigned* pk, unsigned* iso, int64_t nw, hipStream_t s) {
    if (lc < 0 || lt < 0 || lc > kPkHalo || lt > kPkHalo || gc < 0 || gt < 0) return hipErrorInvalidValue;
    const dim3 g1((unsigned)((W + kPkT - 1) / kPkT), (unsigned)((C + kPkC - 1) / kPkC));
    hipLaunchKernelGGL(ispeak_kernel, g1, dim3(kPkT), 0, s, X, ldx, C, W, lc, lt, th, pk, nw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t n = (int64_t)C * nw;
    hipLaunchKernelGGL(isolated_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pk, C, W, nw, gc, gt, edge,
                       iso);
    return hipGetLastError();
}

int snippet_scan_blocks(int64_t nwords) { return (int)((nwords + kScanB - 1) / kScanB); }

hipError_t launch_emit_snippets(const unsigned* iso, int64_t nwords, int64_t nw, int* bsum, long long* total,
                                const float* X, int64_t ldx, int nt, int twav_min, float* clips, int64_t max_clips,
                                int* chan, int* tim, hipStream_t s) {
    const int nb = snippet_scan_blocks(nwords);
    hipLaunchKernelGGL(popcount_blocks_kernel, dim3(nb), dim3(kScanB), 0, s, iso, nwords, bsum);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(kScanB), 0, s, bsum, nb, total);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(emit_snippets_kernel, dim3(nb), dim3(kScanB), 0, s, iso, nwords, nw, bsum, X, ldx, nt, twav_min,
                       clips, max_clips, chan, tim);
    return hipGetLastError();
}

// Q[m][i] = clip_i[m] / ||clip_i|| in FP64 (the row normalisation of spikedetect.py:70), one thread
// per snippet; zero snippets stay zero.
__global__ __launch_bounds__(256) void normalize_clips_kernel(const float* __restrict__ clips, int64_t n, int nt,
                                                              double* __restrict__ Q) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float* c = clips + i * nt;
    double s2 = 0.0;
    for (int m = 0; m < nt; ++m) s2 = fma((double)c[m], (double)c[m], s2);
    const double inv = s2 > 0.0 ? 1.0 / sqrt(s2) : 0.0;
    for (int m = 0; m < nt; ++m) Q[(int64_t)m * n + i] = (double)c[m] * inv;
}

hipError_t launch_normalize_clips(const float* clips, int64_t n, int nt, double* Q, hipStream_t s) {
    hipLaunchKernelGGL(normalize_clips_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, clips, n, nt, Q);
    return hipGetLastError();
}

}  // namespace lss
