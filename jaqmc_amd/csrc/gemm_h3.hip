// Instantiation of the GEMM kernels for PREC_FP16X3 (one translation unit per
// arithmetic mode keeps the builds parallel).  Only launches on tangent rows
// reach it: value-only launches and the batched attention GEMMs of this mode
// run the BF16X6 kernels (gemm.hip).
#include "gemm_impl.h"

int gemm_dispatch_h3(const GemmArgs& a, hipStream_t s) { return dispatch<PREC_FP16X3>(a, s); }
