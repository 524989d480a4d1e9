// Launch-class timing with HIP events, recorded on the launching stream.
#pragma once
#include <hip/hip_runtime.h>

enum ProfClass {
  PROF_GEMM_LAYER_LAP = 0,  // one-electron layer GEMMs, Laplacian mode
  PROF_DET_LAP = 1,         // determinant + Laplacian kernel
  PROF_GEMM_ORB_LAP = 2,    // orbital projection GEMMs, Laplacian mode
  PROF_PAIR_LAP = 3,        // two-electron stream layers, Laplacian mode
  PROF_FORWARD = 4,         // one whole forward (log|psi|) chunk
  PROF_LOCAL_ENERGY = 5,    // one whole local-energy chunk
  PROF_ATTN_LAP = 6,        // Psiformer attention kernels, Laplacian mode
  PROF_NCLASS = 8
};

void prof_begin(int cls, hipStream_t s);
void prof_end(int cls, hipStream_t s);

struct ProfScope {
  int c;
  hipStream_t s;
  ProfScope(int c_, hipStream_t s_) : c(c_), s(s_) { prof_begin(c, s); }
  ~ProfScope() { prof_end(c, s); }
};
