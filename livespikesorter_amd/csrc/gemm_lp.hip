// Reduced-precision variants of the window contractions (placeholder until the MFMA kernels land).
#include "lss_kernels.h"

namespace lss {

hipError_t launch_gemm_lp(int prec, const float* A, int64_t lda, int a_km, const float* B,
                          int64_t ldb, float* C, int64_t ldc, int M, int N, int K, hipStream_t s) {
    (void)prec; (void)A; (void)lda; (void)a_km; (void)B; (void)ldb; (void)C; (void)ldc;
    (void)M; (void)N; (void)K; (void)s;
    return hipErrorNotSupported;
}

}  // namespace lss
