// Shared definitions for the gfx950 kernels of the VMC hot path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16;
typedef _Float16 f16;
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

#define JQ_WAVE 64

// Tangent-row convention used by every Laplacian-mode buffer: for each unit
// (an electron's feature vector, an orbital-matrix element, ...) the rows are
//   row 0        value
//   row 1        Laplacian over all 3N electron coordinates
//   row 2 + c    d/dx_c, c = 3*electron + axis
// padded with zero rows up to a multiple of 16 ("Rp").
#define ROW_VAL 0
#define ROW_LAP 1
#define ROW_J0 2

// Unit (spin-blocked electron) ordering of all per-electron buffers in a
// walker chunk of Bc walkers: all spin-up electrons of all walkers first
// (walker-major), then all spin-down electrons.  A GEMM tile or a spin-channel
// weight choice therefore never straddles channels.
__host__ __device__ inline int unit_of(int w, int i, int Bc, int n_up, int n_dn) {
  return i < n_up ? w * n_up + i : Bc * n_up + w * n_dn + (i - n_up);
}
__host__ __device__ inline int walker_of_unit(int u, int Bc, int n_up, int n_dn) {
  const int uu = Bc * n_up;
  if (u < uu) return u / n_up;  // n_up > 0 whenever spin-up units exist
  return (u - uu) / (n_dn > 0 ? n_dn : 1);
}
__host__ __device__ inline int elec_of_unit(int u, int Bc, int n_up, int n_dn) {
  const int uu = Bc * n_up;
  if (u < uu) return u % n_up;
  return n_up + (u - uu) % (n_dn > 0 ? n_dn : 1);
}

template <typename T> __device__ inline float to_f(T x);
template <> __device__ inline float to_f<float>(float x) { return x; }
template <> __device__ inline float to_f<bf16>(bf16 x) { return (float)x; }
template <typename T> __device__ inline T from_f(float x);
template <> __device__ inline float from_f<float>(float x) { return x; }
template <> __device__ inline bf16 from_f<bf16>(float x) { return (bf16)x; }

// Workgroup -> (row tile, column block).  Workgroups are dispatched to the
// eight XCDs round-robin (workgroup b runs on XCD b % 8, each XCD with its own
// L2), so the column blocks of one row tile go to b, b + 8, b + 16, ...: they
// run back to back on one XCD and the row tile's shared operand is read from
// HBM once instead of once per column block.  Row tiles past the end return
// at once.
constexpr int kXcds = 8;
__host__ __device__ inline void xcd_tile(int b, int ncb, int* rt, int* cb) {
  const int x = b % kXcds, slot = b / kXcds;
  *cb = slot % ncb;
  *rt = (slot / ncb) * kXcds + x;
}
inline int xcd_grid(long nrt, int ncb) { return (int)((nrt + kXcds - 1) / kXcds) * kXcds * ncb; }

// 64-bit DPP move (both halves with the same control)
template <int CTRL>
__device__ inline double dpp_d(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ inline double readlane_d(double x, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// Wavefront sum of a double, uniform result: DPP sums within each 16-lane row
// (xor 1, xor 2, half-row mirror, row mirror), then the four row totals.
__device__ inline double wave_sum_dpp_d(double x) {
  x += dpp_d<0xB1>(x);
  x += dpp_d<0x4E>(x);
  x += dpp_d<0x141>(x);
  x += dpp_d<0x140>(x);
  return readlane_d(x, 0) + readlane_d(x, 16) + readlane_d(x, 32) + readlane_d(x, 48);
}

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double wave_max_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// Orders LDS accesses between the lanes of one wavefront.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

static inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
static inline int ceil_div(long x, long m) { return (int)((x + m - 1) / m); }
