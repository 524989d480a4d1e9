// Instantiation of the GEMM kernels for PREC_F32 (one translation unit per
// arithmetic mode keeps the builds parallel).
#include "gemm_impl.h"

int gemm_dispatch_f32(const GemmArgs& a, hipStream_t s) { return dispatch<PREC_F32>(a, s); }
void bgemm_dispatch_f32(const BGemmArgs& a, hipStream_t s) { dispatch_bg<PREC_F32>(a, s); }
