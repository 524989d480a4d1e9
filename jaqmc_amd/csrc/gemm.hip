// MFMA GEMM with the forward-Laplacian activation epilogue (gfx950).
//
// One kernel family serves every dense layer of the hot path: the FermiNet
// one-electron layers (value + Laplacian + 3N Jacobian rows per electron), the
// spin-mean broadcast term, the orbital projection and the Psiformer
// projections.  Both operands are K-contiguous and staged through LDS in
// K-chunks of 32; each of the 4 waves owns NBW 16-column blocks of the
// workgroup tile and all of its rows, so the cross-row reduction sum_c J_c^2
// that the tanh rule needs stays inside one wavefront (two xor-shuffles).
//
// K-to-lane mapping: within a 32-wide K chunk, lane group g = lane/16 owns
// k = 8g .. 8g+7.  For the BF16 16x16x32 MFMA this is one 16-byte LDS read per
// operand; for the FP32 16x16x4 MFMA, step kk uses k = 8g + kk, so each lane
// again reads its 8 operands with two 16-byte LDS reads.  Row stride in LDS is
// 36 floats / 40 bf16, which keeps 16-byte reads of 8 consecutive rows on
// distinct banks.

#include <type_traits>

#include "kernels.h"

namespace {

template <int PREC> struct ATypeOf { using T = float; };
template <> struct ATypeOf<PREC_BF16> { using T = bf16; };

__device__ inline void split_bf16(float x, bf16& hi, bf16& lo) {
  hi = (bf16)x;
  lo = (bf16)(x - (float)hi);
}

template <int MB, int NBW, int PREC, bool LAP, typename TY>
__global__ __launch_bounds__(256) void gemm_kernel(const GemmArgs a) {
  constexpr int BM = 16 * MB;
  constexpr int BN = 64 * NBW;
  constexpr int KC = 32;
  constexpr bool F32 = PREC == PREC_F32;
  constexpr int LS = F32 ? 36 : 40;  // LDS row stride (elements)
  constexpr int NS = PREC == PREC_BF16X3 ? 2 : 1;
  using TL = typename std::conditional<F32, float, bf16>::type;
  __shared__ __attribute__((aligned(16))) TL As[NS][BM * LS];
  __shared__ __attribute__((aligned(16))) TL Bs[NS][BN * LS];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wcol = wave * NBW * 16;
  const long row0 = (long)blockIdx.x * BM;
  const int col0 = blockIdx.y * BN;

  f32x4 acc[MB][NBW];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NBW; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < a.K; k0 += KC) {
    // ---- global -> LDS ----
    if constexpr (PREC == PREC_BF16) {
      const bf16* X = (const bf16*)a.X;
      const bf16* W = (const bf16*)a.Wt;
      for (int idx = tid; idx < BM * 4; idx += 256) {
        const int r = idx >> 2, q = idx & 3;
        const long gr = row0 + r;
        uint4 v = {0u, 0u, 0u, 0u};
        if (gr < a.rows) v = *(const uint4*)(X + gr * a.ldx + k0 + q * 8);
        *(uint4*)(&As[0][r * LS + q * 8]) = v;
      }
      for (int idx = tid; idx < BN * 4; idx += 256) {
        const int c = idx >> 2, q = idx & 3;
        const int gc = col0 + c;
        uint4 v = {0u, 0u, 0u, 0u};
        if (gc < a.ncols) v = *(const uint4*)(W + (long)gc * a.ldw + k0 + q * 8);
        *(uint4*)(&Bs[0][c * LS + q * 8]) = v;
      }
    } else {
      const float* X = (const float*)a.X;
      for (int idx = tid; idx < BM * 8; idx += 256) {
        const int r = idx >> 3, q = idx & 7;
        const long gr = row0 + r;
        float4 v = {0.f, 0.f, 0.f, 0.f};
        if (gr < a.rows) v = *(const float4*)(X + gr * a.ldx + k0 + q * 4);
        if constexpr (F32) {
          *(float4*)(&As[0][r * LS + q * 4]) = v;
        } else {
          bf16 h[4], l[4];
          split_bf16(v.x, h[0], l[0]);
          split_bf16(v.y, h[1], l[1]);
          split_bf16(v.z, h[2], l[2]);
          split_bf16(v.w, h[3], l[3]);
          *(uint2*)(&As[0][r * LS + q * 4]) = *(uint2*)h;
          *(uint2*)(&As[1][r * LS + q * 4]) = *(uint2*)l;
        }
      }
      if constexpr (F32) {
        const float* W = (const float*)a.Wt;
        for (int idx = tid; idx < BN * 8; idx += 256) {
          const int c = idx >> 3, q = idx & 7;
          const int gc = col0 + c;
          float4 v = {0.f, 0.f, 0.f, 0.f};
          if (gc < a.ncols) v = *(const float4*)(W + (long)gc * a.ldw + k0 + q * 4);
          *(float4*)(&Bs[0][c * LS + q * 4]) = v;
        }
      } else {
        const bf16* Wh = (const bf16*)a.Wt;
        const bf16* Wl = (const bf16*)a.Wt_lo;
        for (int idx = tid; idx < BN * 4; idx += 256) {
          const int c = idx >> 2, q = idx & 3;
          const int gc = col0 + c;
          uint4 vh = {0u, 0u, 0u, 0u}, vl = {0u, 0u, 0u, 0u};
          if (gc < a.ncols) {
            vh = *(const uint4*)(Wh + (long)gc * a.ldw + k0 + q * 8);
            vl = *(const uint4*)(Wl + (long)gc * a.ldw + k0 + q * 8);
          }
          *(uint4*)(&Bs[0][c * LS + q * 8]) = vh;
          *(uint4*)(&Bs[1][c * LS + q * 8]) = vl;
        }
      }
    }
    __syncthreads();

    // ---- MFMA ----
    if constexpr (F32) {
      float bf[NBW][8];
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb) {
        const float* p = &Bs[0][(wcol + nb * 16 + lr) * LS + lg * 8];
        *(float4*)&bf[nb][0] = *(const float4*)p;
        *(float4*)&bf[nb][4] = *(const float4*)(p + 4);
      }
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        float af[8];
        const float* p = &As[0][(mb * 16 + lr) * LS + lg * 8];
        *(float4*)&af[0] = *(const float4*)p;
        *(float4*)&af[4] = *(const float4*)(p + 4);
#pragma unroll
        for (int kk = 0; kk < 8; ++kk)
#pragma unroll
          for (int nb = 0; nb < NBW; ++nb)
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[kk], bf[nb][kk], acc[mb][nb], 0, 0, 0);
      }
    } else {
      bf16x8 bh[NBW], bl[NBW];
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb) {
        bh[nb] = *(const bf16x8*)&Bs[0][(wcol + nb * 16 + lr) * LS + lg * 8];
        if constexpr (NS == 2) bl[nb] = *(const bf16x8*)&Bs[1][(wcol + nb * 16 + lr) * LS + lg * 8];
      }
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const bf16x8 ah = *(const bf16x8*)&As[0][(mb * 16 + lr) * LS + lg * 8];
        if constexpr (NS == 2) {
          const bf16x8 al = *(const bf16x8*)&As[1][(mb * 16 + lr) * LS + lg * 8];
#pragma unroll
          for (int nb = 0; nb < NBW; ++nb) {
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[nb], acc[mb][nb], 0, 0, 0);
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[nb], acc[mb][nb], 0, 0, 0);
          }
        }
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb)
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[nb], acc[mb][nb], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  // ---- epilogue ----
  // acc[mb][nb][v] holds row mb*16 + 4*lg + v, column col0 + wcol + nb*16 + lr.
  TY* Y = (TY*)a.Y;
  const TY* Res = (const TY*)a.Res;
  if constexpr (LAP) {
    const long unit = blockIdx.x;
    const int w = a.G ? walker_of_unit((int)unit, a.Bc, a.n_up, a.n_dn) : 0;
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      const int n = col0 + wcol + nb * 16 + lr;
      const bool ok = n < a.ncols;
      float y[MB][4];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = mb * 16 + lg * 4 + v;
          float t = acc[mb][nb][v];
          if (ok && a.G) t += a.G[((long)w * BM + r) * a.ldg + n];
          if (ok && r == ROW_VAL && a.bias) t += a.bias[n];
          y[mb][v] = t;
        }
      if (a.act == ACT_TANH) {
        float sq = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (mb * 16 + lg * 4 + v >= ROW_J0) sq += y[mb][v] * y[mb][v];
        sq += __shfl_xor(sq, 16);
        sq += __shfl_xor(sq, 32);
        const float x = __shfl(y[0][0], lr);
        const float l = __shfl(y[0][1], lr);
        const float t = tanhf(x);
        const float d1 = 1.f - t * t;
        const float d2 = -2.f * t * d1;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = mb * 16 + lg * 4 + v;
            y[mb][v] = r == ROW_VAL ? t : (r == ROW_LAP ? d1 * l + d2 * sq : d1 * y[mb][v]);
          }
      }
      if (ok) {
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const long row = row0 + mb * 16 + lg * 4 + v;
            float o = y[mb][v];
            if (Res) o = (o + to_f<TY>(Res[row * a.ldres + n])) * a.res_scale;
            Y[row * a.ldy + n] = from_f<TY>(o);
          }
      }
    }
  } else {
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      const int n = col0 + wcol + nb * 16 + lr;
      if (n >= a.ncols) continue;
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const long row = row0 + mb * 16 + lg * 4 + v;
          if (row >= a.rows) continue;
          float t = acc[mb][nb][v];
          if (a.G) t += a.G[(long)walker_of_unit((int)row, a.Bc, a.n_up, a.n_dn) * a.ldg + n];
          if (a.bias) t += a.bias[n];
          if (a.act == ACT_TANH) t = tanhf(t);
          if (Res) t = (t + to_f<TY>(Res[row * a.ldres + n])) * a.res_scale;
          Y[row * a.ldy + n] = from_f<TY>(t);
        }
    }
  }
}

template <int MB, bool LAP, int PREC, typename TY>
void launch_one(const GemmArgs& a, hipStream_t s) {
  constexpr int BM = 16 * MB, BN = 128;
  dim3 grid(LAP ? a.rows / a.R : ceil_div(a.rows, BM), ceil_div(a.ncols, BN));
  hipLaunchKernelGGL((gemm_kernel<MB, 2, PREC, LAP, TY>), grid, dim3(256), 0, s, a);
}

template <int PREC, typename TY>
int dispatch(const GemmArgs& a, hipStream_t s) {
  if (a.R <= 1) {
    launch_one<4, false, PREC, TY>(a, s);
    return 0;
  }
  switch (a.R) {
    case 16: launch_one<1, true, PREC, TY>(a, s); return 0;
    case 32: launch_one<2, true, PREC, TY>(a, s); return 0;
    case 48: launch_one<3, true, PREC, TY>(a, s); return 0;
    case 64: launch_one<4, true, PREC, TY>(a, s); return 0;
    case 96: launch_one<6, true, PREC, TY>(a, s); return 0;
    case 128: launch_one<8, true, PREC, TY>(a, s); return 0;
    case 192: launch_one<12, true, PREC, TY>(a, s); return 0;
    default: return -1;
  }
}

}  // namespace

// Supported Laplacian-mode unit heights (rows per unit after padding).
int gemm_unit_rows(int R) {
  const int opts[] = {16, 32, 48, 64, 96, 128, 192};
  for (int o : opts)
    if (R <= o) return o;
  return -1;
}

int launch_gemm(const GemmArgs& a, int prec, bool y_bf16, hipStream_t s) {
  if (a.K % 32 != 0 || a.ldx % 8 != 0 || a.ldw % 8 != 0) return -1;
  if (a.R > 1 && a.rows % a.R != 0) return -1;
  if (a.rows == 0) return 0;
  switch (prec) {
    case PREC_F32: return y_bf16 ? -1 : dispatch<PREC_F32, float>(a, s);
    case PREC_BF16: return y_bf16 ? dispatch<PREC_BF16, bf16>(a, s) : dispatch<PREC_BF16, float>(a, s);
    case PREC_BF16X3: return y_bf16 ? -1 : dispatch<PREC_BF16X3, float>(a, s);
  }
  return -1;
}
