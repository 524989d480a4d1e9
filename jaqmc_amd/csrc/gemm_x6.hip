// Instantiation of the GEMM kernels for PREC_BF16X6 (one translation unit per
// arithmetic mode keeps the builds parallel).
#include "gemm_impl.h"

int gemm_dispatch_x6(const GemmArgs& a, hipStream_t s) { return dispatch<PREC_BF16X6>(a, s); }
void bgemm_dispatch_x6(const BGemmArgs& a, hipStream_t s) { dispatch_bg<PREC_BF16X6>(a, s); }
