// MFMA GEMM with the forward-Laplacian activation epilogue (gfx950).
//
// One kernel family serves every dense layer of the hot path: the FermiNet
// one-electron layers (value + Laplacian + 3N Jacobian rows per electron), the
// spin-mean broadcast term, the orbital projection and the Psiformer
// projections.  Both operands are K-contiguous FP32 in HBM and staged through
// LDS in K-chunks of 32, with the next chunk prefetched into registers while
// the current one is multiplied.  Each of the 4 waves owns NBW 16-column blocks
// of the tile and *all* of its rows, so the cross-row reduction sum_c J_c^2 of
// the tanh rule stays inside one wavefront (two xor-shuffles).
//
// Arithmetic:
//   PREC_F32     v_mfma_f32_16x16x4_f32
//   PREC_BF16    one BF16 term per operand, v_mfma_f32_16x16x32_bf16
//   PREC_BF16X3  x = h + l (two BF16 terms): hh + hl + lh
//   PREC_BF16X6  x = h + m + l (three BF16 terms): hh + hm + mh + hl + lh + mm,
//                products down to 2^-16 relative are kept, so the result has
//                FP32-level error (~2^-24) with FP32 accumulation.
// Activations are split into BF16 terms when staged into LDS; weights are
// constant and arrive pre-split as NS BF16 planes (GemmArgs::Wt_lo).
//
// K-to-lane mapping: within a 32-wide K chunk, lane group g = lane/16 owns
// k = 8g .. 8g+7.  For the BF16 16x16x32 MFMA this is one 16-byte LDS read per
// operand term; for the FP32 16x16x4 MFMA, step kk uses k = 8g + kk, so each
// lane reads its 8 operands with two 16-byte LDS reads.  LDS row stride is 36
// floats / 40 bf16, which keeps 16-byte reads of 8 consecutive rows on
// distinct banks.

#include <type_traits>

#include "kernels.h"

namespace {

template <int PREC> struct NSplit { static constexpr int v = 1; };
template <> struct NSplit<PREC_BF16X3> { static constexpr int v = 2; };
template <> struct NSplit<PREC_BF16X6> { static constexpr int v = 3; };

__device__ inline void split3(float x, bf16* t, int ns) {
  float r = x;
  for (int i = 0; i < ns; ++i) {
    const bf16 h = (bf16)r;
    t[i] = h;
    r = r - (float)h;
  }
}

// MBU: 16-row blocks per unit (Laplacian mode); UPT: units per tile; NBW:
// 16-column blocks per wave.  BM = 16*MBU*UPT, BN = 64*NBW.
template <int MBU, int UPT, int NBW, int PREC, bool LAP>
__global__ __launch_bounds__(256) void gemm_kernel(const GemmArgs a) {
  constexpr int MB = MBU * UPT;
  constexpr int BM = 16 * MB;
  constexpr int BN = 64 * NBW;
  constexpr int KC = 32;
  constexpr bool F32 = PREC == PREC_F32;
  constexpr int NS = NSplit<PREC>::v;
  constexpr int LS = F32 ? 36 : 40;  // LDS row stride (elements)
  using TL = typename std::conditional<F32, float, bf16>::type;
  constexpr int NSL = F32 ? 1 : NS;
  // One LDS arena: operand tiles during the K loop, the output tile (FP32,
  // row stride BN + 4) in the epilogue.
  constexpr int MAIN_BYTES = NSL * (BM + BN) * LS * (int)sizeof(TL);
  constexpr int SLD = BN + 4;
  constexpr int EPI_BYTES = BM * SLD * 4 + UPT * BN * 4;
  constexpr int ARENA = MAIN_BYTES > EPI_BYTES ? MAIN_BYTES : EPI_BYTES;
  __shared__ __attribute__((aligned(16))) char arena[ARENA];
  TL (*As)[BM * LS] = (TL (*)[BM * LS])arena;
  TL (*Bs)[BN * LS] = (TL (*)[BN * LS])(arena + NSL * BM * LS * sizeof(TL));

  // register prefetch: A and B chunks as float4 (8 per 32-wide row)
  constexpr int AV = (BM * 8 + 255) / 256;
  constexpr int BV = F32 ? (BN * 8 + 255) / 256 : 1;
  constexpr int BVB = F32 ? 1 : (BN * 4 + 255) / 256;
  float4 pa[AV], pb[BV];
  uint4 pbb[NS][BVB];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wcol = wave * NBW * 16;
  const long row0 = (long)blockIdx.x * BM;
  const int col0 = blockIdx.y * BN;
  const float* X = (const float*)a.X + (long)blockIdx.z * a.sxb;
  const float* W = (const float*)a.Wt;

  auto gload = [&](int k0) {
#pragma unroll
    for (int v = 0; v < AV; ++v) {
      const int idx = tid + v * 256;
      const int r = idx >> 3, q = idx & 7;
      const long gr = row0 + r;
      pa[v] = (idx < BM * 8 && gr < a.rows) ? *(const float4*)(X + gr * a.ldx + k0 + q * 4)
                                             : float4{0.f, 0.f, 0.f, 0.f};
    }
    if constexpr (F32) {
#pragma unroll
      for (int v = 0; v < BV; ++v) {
        const int idx = tid + v * 256;
        const int c = idx >> 3, q = idx & 7;
        const int gc = col0 + c;
        pb[v] = (idx < BN * 8 && gc < a.ncols) ? *(const float4*)(W + (long)gc * a.ldw + k0 + q * 4)
                                               : float4{0.f, 0.f, 0.f, 0.f};
      }
    } else {
      const bf16* Wb = (const bf16*)a.Wt_lo;
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int v = 0; v < BVB; ++v) {
          const int idx = tid + v * 256;
          const int c = idx >> 2, q = idx & 3;
          const int gc = col0 + c;
          pbb[s][v] = (idx < BN * 4 && gc < a.ncols)
                          ? *(const uint4*)(Wb + s * a.wplane + (long)gc * a.ldw + k0 + q * 8)
                          : uint4{0u, 0u, 0u, 0u};
        }
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int v = 0; v < AV; ++v) {
      const int idx = tid + v * 256;
      if (idx >= BM * 8) break;
      const int r = idx >> 3, q = idx & 7;
      if constexpr (F32) {
        *(float4*)(&As[0][r * LS + q * 4]) = pa[v];
      } else {
        bf16 t[4][NS];
        split3(pa[v].x, t[0], NS);
        split3(pa[v].y, t[1], NS);
        split3(pa[v].z, t[2], NS);
        split3(pa[v].w, t[3], NS);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          bf16 h[4] = {t[0][s], t[1][s], t[2][s], t[3][s]};
          *(uint2*)(&As[s][r * LS + q * 4]) = *(uint2*)h;
        }
      }
    }
    if constexpr (F32) {
#pragma unroll
      for (int v = 0; v < BV; ++v) {
        const int idx = tid + v * 256;
        if (idx >= BN * 8) break;
        const int c = idx >> 3, q = idx & 7;
        *(float4*)(&Bs[0][c * LS + q * 4]) = pb[v];
      }
    } else {
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int v = 0; v < BVB; ++v) {
          const int idx = tid + v * 256;
          if (idx >= BN * 4) break;
          const int c = idx >> 2, q = idx & 3;
          *(uint4*)(&Bs[s][c * LS + q * 8]) = pbb[s][v];
        }
    }
  };

  f32x4 acc[MB][NBW];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NBW; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  gload(0);
  for (int k0 = 0; k0 < a.K; k0 += KC) {
    lstore();
    __syncthreads();
    if (k0 + KC < a.K) gload(k0 + KC);

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
      bf16x8 bt[NS][NBW];
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) bt[s][nb] = *(const bf16x8*)&Bs[s][(wcol + nb * 16 + lr) * LS + lg * 8];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        bf16x8 at[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) at[s] = *(const bf16x8*)&As[s][(mb * 16 + lr) * LS + lg * 8];
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) {
          // small terms first
          if constexpr (NS == 3) {
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[1], bt[1][nb], acc[mb][nb], 0, 0, 0);
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[2][nb], acc[mb][nb], 0, 0, 0);
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[2], bt[0][nb], acc[mb][nb], 0, 0, 0);
          }
          if constexpr (NS >= 2) {
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[1][nb], acc[mb][nb], 0, 0, 0);
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[1], bt[0][nb], acc[mb][nb], 0, 0, 0);
          }
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[0][nb], acc[mb][nb], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }

  // ---- epilogue ----
  // Stage the accumulators in LDS, then make every global access a coalesced
  // 16-byte one: y = acc + G[walker, r] + bias[value rows]; in Laplacian mode
  // one thread per (unit, column) applies the tanh rule down the unit's rows;
  // then out = f(y) (+ residual) is stored row-major.
  // acc[mb][nb][v] holds row mb*16 + 4*lg + v, column wcol + nb*16 + lr.
  float* St = (float*)arena;
  float* D1 = St + BM * SLD;  // (UPT, BN) tanh' per unit column
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb)
#pragma unroll
      for (int v = 0; v < 4; ++v) St[(mb * 16 + lg * 4 + v) * SLD + wcol + nb * 16 + lr] = acc[mb][nb][v];
  __syncthreads();
  constexpr int RU = LAP ? 16 * MBU : 1;
  constexpr int C4 = BN / 4;
  // pass 1: broadcast term and bias (pre-activation)
  const bool vec_g = !a.G || (a.ldg % 4 == 0);
  if (a.G || a.bias) {
    for (int idx = tid; idx < BM * C4; idx += 256) {
      const int r = idx / C4, c = (idx % C4) * 4;
      const long grow = row0 + r;
      const int n = col0 + c;
      if (grow >= a.rows || n >= a.ncols) continue;
      const int rr = r % RU;
      const float* gp = nullptr;
      if (a.G) {
        const long unit = LAP ? grow / RU : grow;
        const int w = a.wmaj ? (int)(unit / (a.n_up + a.n_dn)) : walker_of_unit((int)unit, a.Bc, a.n_up, a.n_dn);
        gp = a.G + ((long)w * RU + rr) * a.ldg + n;
      }
      const bool bias_here = a.bias && rr == ROW_VAL;
      if (n + 4 <= a.ncols && vec_g) {
        float4 y = *(float4*)&St[r * SLD + c];
        if (gp) {
          const float4 g = *(const float4*)gp;
          y.x += g.x;
          y.y += g.y;
          y.z += g.z;
          y.w += g.w;
        }
        if (bias_here) {
          y.x += a.bias[n];
          y.y += a.bias[n + 1];
          y.z += a.bias[n + 2];
          y.w += a.bias[n + 3];
        }
        *(float4*)&St[r * SLD + c] = y;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (n + q >= a.ncols) continue;
          float y = St[r * SLD + c + q];
          if (gp) y += gp[q];
          if (bias_here) y += a.bias[n + q];
          St[r * SLD + c + q] = y;
        }
      }
    }
    __syncthreads();
  }
  // pass 2 (Laplacian mode, tanh): per (unit, column) rule down the rows
  if (LAP && a.act == ACT_TANH) {
    for (int idx = tid; idx < UPT * BN; idx += 256) {
      const int u = idx / BN, c = idx % BN;
      float* col = St + u * RU * SLD + c;
      float sq = 0.f;
      for (int rr = ROW_J0; rr < RU; ++rr) {
        const float j = col[rr * SLD];
        sq += j * j;
      }
      const float t = tanhf(col[0]);
      const float d1 = 1.f - t * t;
      col[ROW_LAP * SLD] = d1 * col[ROW_LAP * SLD] - 2.f * t * d1 * sq;
      col[0] = t;
      D1[u * BN + c] = d1;
    }
    __syncthreads();
  }
  // pass 3: activation of the remaining rows, residual, coalesced store
  float* Y = (float*)a.Y + (long)blockIdx.z * a.syb;
  const float* Res = (const float*)a.Res;
  const bool vec_io = (a.ldy % 4 == 0) && (!Res || a.ldres % 4 == 0) && (a.hs_H == 0 || a.hs_dh % 4 == 0);
  for (int idx = tid; idx < BM * C4; idx += 256) {
    const int r = idx / C4, c = (idx % C4) * 4;
    const long grow = row0 + r;
    const int n = col0 + c;
    if (grow >= a.rows || n >= a.ncols) continue;
    float y[4], d[4] = {1.f, 1.f, 1.f, 1.f};
    *(float4*)y = *(float4*)&St[r * SLD + c];
    const bool tail = !(n + 4 <= a.ncols && vec_io);
    if (LAP && a.act == ACT_TANH && (r % RU) >= ROW_J0) *(float4*)d = *(float4*)&D1[(r / RU) * BN + c];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (a.act == ACT_TANH) y[q] = LAP ? y[q] * d[q] : tanhf(y[q]);
    }
    if (a.hs_H > 0) {
      // head-split destination (Psiformer Q/K/V), walker-major units
      const long unit = LAP ? grow / RU : grow;
      const int rr = LAP ? (int)(grow % RU) : 0;
      const int w = (int)(unit / a.hs_N), i = (int)(unit % a.hs_N);
      const int part = n / a.hs_D, hh = (n % a.hs_D) / a.hs_dh, dd = n % a.hs_dh;
      float* dst = Y + ((((long)w * 3 + part) * a.hs_H + hh) * a.hs_N + i) * (long)RU * a.hs_dh + (long)rr * a.hs_dh + dd;
      if (!tail) {
        *(float4*)dst = *(float4*)y;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (n + q < a.ncols) dst[q] = y[q];
      }
      continue;
    }
    if (!tail) {
      if (Res) {
        const float4 rq = *(const float4*)(Res + grow * a.ldres + n);
        y[0] = (y[0] + rq.x) * a.res_scale;
        y[1] = (y[1] + rq.y) * a.res_scale;
        y[2] = (y[2] + rq.z) * a.res_scale;
        y[3] = (y[3] + rq.w) * a.res_scale;
      }
      *(float4*)(Y + grow * a.ldy + n) = *(float4*)y;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (n + q >= a.ncols) continue;
        float o = y[q];
        if (Res) o = (o + Res[grow * a.ldres + n + q]) * a.res_scale;
        Y[grow * a.ldy + n + q] = o;
      }
    }
  }
}

template <int MBU, int UPT, int NBW, int PREC, bool LAP>
void launch_one(const GemmArgs& a, hipStream_t s) {
  constexpr int BM = 16 * MBU * UPT, BN = 64 * NBW;
  dim3 grid(ceil_div(a.rows, BM), ceil_div(a.ncols, BN), a.nbatch > 1 ? a.nbatch : 1);
  hipLaunchKernelGGL((gemm_kernel<MBU, UPT, NBW, PREC, LAP>), grid, dim3(256), 0, s, a);
}

int g_gemm_variant = 0;  // 0: default table; else UPT*10 + NBW (benchmark override)

template <int MBU, int PREC>
bool launch_lap(const GemmArgs& a, int upt, int nbw, hipStream_t s) {
#define JQ_TRY(U, NB)                                                           \
  if constexpr (MBU * U <= 12 && MBU * U * NB <= 36) {                          \
    if (upt == U && nbw == NB) {                                                \
      launch_one<MBU, U, NB, PREC, true>(a, s);                                 \
      return true;                                                              \
    }                                                                           \
  }
  JQ_TRY(1, 2)
  JQ_TRY(1, 4)
  JQ_TRY(2, 2)
  JQ_TRY(2, 4)
  JQ_TRY(4, 2)
  JQ_TRY(4, 4)
#undef JQ_TRY
  return false;
}

// Default (UPT, NBW) per unit height and precision, from tile sweeps on
// MI355X (K = 256-320, 256 output columns).
void default_cfg(int R, int prec, bool wide, int* upt, int* nbw) {
  (void)prec;
  (void)wide;
  *nbw = 2;
  switch (R) {
    case 16: *upt = 4; break;
    case 32: *upt = 2; break;
    case 48: *upt = 2; break;
    case 64: *upt = 2; break;
    default: *upt = 1; break;
  }
}

template <int PREC>
int dispatch(const GemmArgs& a, hipStream_t s) {
  const bool wide = a.ncols > 128;
  if (a.R <= 1) {
    if (wide)
      launch_one<8, 1, 4, PREC, false>(a, s);
    else
      launch_one<8, 1, 2, PREC, false>(a, s);
    return 0;
  }
  int upt, nbw;
  default_cfg(a.R, PREC, wide, &upt, &nbw);
  if (g_gemm_variant > 0) {
    upt = g_gemm_variant / 10;
    nbw = g_gemm_variant % 10;
  }
  bool ok = false;
  switch (a.R) {
    case 16: ok = launch_lap<1, PREC>(a, upt, nbw, s); break;
    case 32: ok = launch_lap<2, PREC>(a, upt, nbw, s); break;
    case 48: ok = launch_lap<3, PREC>(a, upt, nbw, s); break;
    case 64: ok = launch_lap<4, PREC>(a, upt, nbw, s); break;
    case 96: ok = launch_lap<6, PREC>(a, upt, nbw, s); break;
    case 128: ok = launch_lap<8, PREC>(a, upt, nbw, s); break;
    case 192: ok = launch_lap<12, PREC>(a, upt, nbw, s); break;
    default: break;
  }
  return ok ? 0 : -1;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched GEMM on activation operands (attention contractions).  Same K-chunk
// pipeline and MFMA arithmetic as gemm_kernel; both operands are FP32 in HBM
// and split into BF16 terms on the way into LDS for the BF16 modes.
// ---------------------------------------------------------------------------
template <int MB, int NBW, int PREC, bool NN>
__global__ __launch_bounds__(256) void bgemm_kernel(const BGemmArgs a) {
  constexpr int BM = 16 * MB;
  constexpr int BN = 64 * NBW;
  constexpr int KC = 32;
  constexpr bool F32 = PREC == PREC_F32;
  constexpr int NS = NSplit<PREC>::v;
  constexpr int LS = F32 ? 36 : 40;
  using TL = typename std::conditional<F32, float, bf16>::type;
  constexpr int NSL = F32 ? 1 : NS;
  __shared__ __attribute__((aligned(16))) TL As[NSL][BM * LS];
  __shared__ __attribute__((aligned(16))) TL Bs[NSL][BN * LS];
  constexpr int AV = (BM * 8 + 255) / 256;
  constexpr int BV = (BN * 8 + 255) / 256;
  float4 pa[AV], pb[BV];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wcol = wave * NBW * 16;
  const long row0 = (long)blockIdx.x * BM;
  const int col0 = blockIdx.y * BN;
  const int z = blockIdx.z;
  const int bw = z / a.nh, bh = z % a.nh;
  const float* X = a.X + bw * a.sxw + bh * a.sxh;
  const float* W = a.W + bw * a.sww + bh * a.swh;

  // Both tiles are staged as rows of 32 K-values: A row r, B row n (output
  // column).  NT: a thread loads 4 consecutive k of one row (float4).  NN: a
  // thread gathers 4 consecutive k of one column n (4 coalesced scalar loads
  // across the wave).  Each thread then writes its 4 values with one packed
  // store per BF16 term (or one float4).
  auto gload = [&](int k0) {
#pragma unroll
    for (int v = 0; v < AV; ++v) {
      const int idx = tid + v * 256;
      const int r = idx >> 3, q = idx & 7;
      const long gr = row0 + r;
      pa[v] = (idx < BM * 8 && gr < a.rows) ? *(const float4*)(X + gr * a.ldx + k0 + q * 4)
                                             : float4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int v = 0; v < BV; ++v) {
      const int idx = tid + v * 256;
      if constexpr (!NN) {
        const int c = idx >> 3, q = idx & 7;
        const int gc = col0 + c;
        pb[v] = (idx < BN * 8 && gc < a.ncols) ? *(const float4*)(W + (long)gc * a.ldw + k0 + q * 4)
                                               : float4{0.f, 0.f, 0.f, 0.f};
      } else {
        // idx -> (k quad q = idx / BN, column c = idx % BN): lanes walk columns
        const int q = idx / BN, c = idx % BN;
        const int gc = col0 + c;
        float e[4] = {0.f, 0.f, 0.f, 0.f};
        if (idx < BN * 8 && gc < a.ncols) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int gk = k0 + 4 * q + u;
            if (gk < a.kvalid) e[u] = W[(long)gk * a.ldw + gc];
          }
        }
        pb[v] = float4{e[0], e[1], e[2], e[3]};
      }
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int v = 0; v < AV; ++v) {
      const int idx = tid + v * 256;
      if (idx >= BM * 8) break;
      const int r = idx >> 3, q = idx & 7;
      if constexpr (F32) {
        *(float4*)(&As[0][r * LS + q * 4]) = pa[v];
      } else {
        bf16 t[4][NS];
        split3(pa[v].x, t[0], NS);
        split3(pa[v].y, t[1], NS);
        split3(pa[v].z, t[2], NS);
        split3(pa[v].w, t[3], NS);
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) {
          bf16 h[4] = {t[0][s2], t[1][s2], t[2][s2], t[3][s2]};
          *(uint2*)(&As[s2][r * LS + q * 4]) = *(uint2*)h;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < BV; ++v) {
      const int idx = tid + v * 256;
      if (idx >= BN * 8) break;
      int c, q;
      if constexpr (!NN) {
        c = idx >> 3;
        q = idx & 7;
      } else {
        q = idx / BN;
        c = idx % BN;
      }
      if constexpr (F32) {
        *(float4*)(&Bs[0][c * LS + q * 4]) = pb[v];
      } else {
        bf16 t[4][NS];
        split3(pb[v].x, t[0], NS);
        split3(pb[v].y, t[1], NS);
        split3(pb[v].z, t[2], NS);
        split3(pb[v].w, t[3], NS);
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) {
          bf16 h[4] = {t[0][s2], t[1][s2], t[2][s2], t[3][s2]};
          *(uint2*)(&Bs[s2][c * LS + q * 4]) = *(uint2*)h;
        }
      }
    }
  };

  f32x4 acc[MB][NBW];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NBW; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  gload(0);
  for (int k0 = 0; k0 < a.K; k0 += KC) {
    lstore();
    __syncthreads();
    if (k0 + KC < a.K) gload(k0 + KC);
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
      bf16x8 bt[NS][NBW];
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2)
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) bt[s2][nb] = *(const bf16x8*)&Bs[s2][(wcol + nb * 16 + lr) * LS + lg * 8];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        bf16x8 at[NS];
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) at[s2] = *(const bf16x8*)&As[s2][(mb * 16 + lr) * LS + lg * 8];
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) {
          if constexpr (NS == 3) {
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[1], bt[1][nb], acc[mb][nb], 0, 0, 0);
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[2][nb], acc[mb][nb], 0, 0, 0);
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[2], bt[0][nb], acc[mb][nb], 0, 0, 0);
          }
          if constexpr (NS >= 2) {
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[1][nb], acc[mb][nb], 0, 0, 0);
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[1], bt[0][nb], acc[mb][nb], 0, 0, 0);
          }
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[0][nb], acc[mb][nb], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }

  float* Y = a.Y + bw * a.syw + bh * a.syh;
  const float* Res = a.Res ? a.Res + bw * a.syw + bh * a.syh : nullptr;
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) {
    const int n = col0 + wcol + nb * 16 + lr;
    if (n >= a.ncols) continue;
    const long cn = a.ycw > 0 ? (long)(n / a.ycw) * a.ycs + n % a.ycw : n;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long row = row0 + mb * 16 + lg * 4 + v;
        if (row >= a.rows) continue;
        float t = acc[mb][nb][v];
        if (Res) t += Res[row * a.ldy + cn];
        Y[row * a.ldy + cn] = t;
      }
  }
}

template <int MB, int NBW, int PREC>
void launch_bg(const BGemmArgs& a, hipStream_t s) {
  dim3 grid(ceil_div(a.rows, 16 * MB), ceil_div(a.ncols, 64 * NBW), a.nw * a.nh);
  if (a.nn)
    hipLaunchKernelGGL((bgemm_kernel<MB, NBW, PREC, true>), grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((bgemm_kernel<MB, NBW, PREC, false>), grid, dim3(256), 0, s, a);
}

template <int PREC>
void dispatch_bg(const BGemmArgs& a, hipStream_t s) {
  if (a.rows <= 48) {
    if (a.ncols > 128)
      launch_bg<3, 4, PREC>(a, s);
    else
      launch_bg<3, 1, PREC>(a, s);
  } else if (a.ncols > 128) {
    launch_bg<4, 4, PREC>(a, s);
  } else if (a.rows <= 64) {
    launch_bg<4, 1, PREC>(a, s);
  } else {
    launch_bg<8, 1, PREC>(a, s);
  }
}

int launch_bgemm(const BGemmArgs& a, int prec, hipStream_t s) {
  if (a.K % 32 != 0 || a.ldx % 4 != 0 || (!a.nn && a.ldw % 4 != 0) || (a.nn && (a.ldw % 4 || a.ncols % 4)))
    return -1;
  if (a.rows <= 0 || a.ncols <= 0 || a.nw * a.nh <= 0) return 0;
  switch (prec) {
    case PREC_F32: dispatch_bg<PREC_F32>(a, s); return 0;
    case PREC_BF16: dispatch_bg<PREC_BF16>(a, s); return 0;
    case PREC_BF16X3: dispatch_bg<PREC_BF16X3>(a, s); return 0;
    case PREC_BF16X6: dispatch_bg<PREC_BF16X6>(a, s); return 0;
  }
  return -1;
}

// Supported Laplacian-mode unit heights (rows per unit after padding).
int gemm_unit_rows(int R) {
  const int opts[] = {16, 32, 48, 64, 96, 128, 192};
  for (int o : opts)
    if (R <= o) return o;
  return -1;
}

int launch_gemm(const GemmArgs& a, int prec, bool y_bf16, hipStream_t s) {
  if (y_bf16) return -1;
  if (a.K % 32 != 0 || a.ldx % 4 != 0 || a.ldw % 4 != 0) return -1;
  if (a.R > 1 && a.rows % a.R != 0) return -1;
  if (a.rows == 0) return 0;
  switch (prec) {
    case PREC_F32: return dispatch<PREC_F32>(a, s);
    case PREC_BF16: return dispatch<PREC_BF16>(a, s);
    case PREC_BF16X3: return dispatch<PREC_BF16X3>(a, s);
    case PREC_BF16X6: return dispatch<PREC_BF16X6>(a, s);
  }
  return -1;
}

// Benchmark hook: time `reps` launches of one GEMM configuration on synthetic
// device buffers (HIP events on the null stream).  variant as g_gemm_variant.
extern "C" int jq_bench_gemm(int prec, int R, int rows, int K, int ncols, int act, int variant, int reps,
                             float* ms_out) {
  if (rows <= 0 || K <= 0 || ncols <= 0 || K % 32 || (R > 1 && rows % R) || reps <= 0) return -1;
  const int ns = prec == PREC_BF16X6 ? 3 : (prec == PREC_BF16X3 ? 2 : 1);
  float *X = nullptr, *W = nullptr, *Y = nullptr, *G = nullptr, *b = nullptr;
  bf16* Wb = nullptr;
  const long nW = (long)ncols * K;
  int rc = 0;
  if (hipMalloc(&X, (size_t)rows * K * 4) || hipMalloc(&W, nW * 4) || hipMalloc(&Y, (size_t)rows * ncols * 4) ||
      hipMalloc(&Wb, nW * ns * 2) || hipMalloc(&G, (size_t)ncols * 4 * (R > 1 ? R : 1) * (rows / (R > 1 ? R : 1) / 14 + 1)) ||
      hipMalloc(&b, ncols * 4))
    rc = -3;
  if (!rc) {
    (void)hipMemset(X, 0, (size_t)rows * K * 4);
    (void)hipMemset(W, 0, nW * 4);
    (void)hipMemset(Wb, 0, nW * ns * 2);
    (void)hipMemset(b, 0, ncols * 4);
    (void)hipMemset(G, 0, (size_t)ncols * 4 * (R > 1 ? R : 1) * (rows / (R > 1 ? R : 1) / 14 + 1));
    GemmArgs a{};
    a.X = X;
    a.ldx = K;
    a.Wt = W;
    a.ldw = K;
    a.Wt_lo = Wb;
    a.wplane = nW;
    a.Y = Y;
    a.ldy = ncols;
    a.rows = rows;
    a.K = K;
    a.ncols = ncols;
    a.bias = b;
    a.R = R;
    a.act = act ? ACT_TANH : ACT_NONE;
    a.res_scale = 1.f;
    if (act == 2 && ncols <= K) {  // realistic layer epilogue: G broadcast + residual
      const int units = R > 1 ? rows / R : rows;
      a.Bc = units / 14 > 0 ? units / 14 : 1;
      a.n_up = 7;
      a.n_dn = 7;
      a.G = G;
      a.ldg = ncols;
      a.Res = X;
      a.ldres = K;
      if ((long)a.Bc * (R > 1 ? R : 1) > 8192L * (R > 1 ? R : 1)) rc = -1;
    }
    const int old = g_gemm_variant;
    g_gemm_variant = variant;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    if (!rc) rc = launch_gemm(a, prec, false, 0);
    (void)hipEventRecord(e0, 0);
    for (int i = 0; i < reps && !rc; ++i) rc = launch_gemm(a, prec, false, 0);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_out = ms / reps;
    g_gemm_variant = old;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(X);
  (void)hipFree(W);
  (void)hipFree(Y);
  (void)hipFree(Wb);
  (void)hipFree(G);
  (void)hipFree(b);
  return rc;
}
