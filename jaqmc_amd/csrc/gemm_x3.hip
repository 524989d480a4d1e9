// Instantiation of the GEMM kernels for PREC_BF16X3 (one translation unit per
// arithmetic mode keeps the builds parallel).
#include "gemm_impl.h"

int gemm_dispatch_x3(const GemmArgs& a, hipStream_t s) { return dispatch<PREC_BF16X3>(a, s); }
void bgemm_dispatch_x3(const BGemmArgs& a, hipStream_t s) { dispatch_bg<PREC_BF16X3>(a, s); }
