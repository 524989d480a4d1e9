// MFMA GEMM kernels with the forward-Laplacian activation epilogue (gfx950).
// Included once per arithmetic mode (gemm_<mode>.hip instantiates it).
//
// One kernel family serves every dense layer of the hot path: the FermiNet
// one-electron layers (value + Laplacian + 3N Jacobian rows per electron), the
// spin-mean broadcast term, the orbital projection and the Psiformer
// projections.  Both operands are K-contiguous FP32 in HBM and staged through
// LDS in K-chunks of 32, with the next chunk prefetched into registers while
// the current one is multiplied (weights: see the K loop).  Each wave (4, or 8 in
// 256-column tiles of two-block waves) owns NBW 16-column blocks
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
//   PREC_FP16X3  x = h + l (two FP16 terms, 11-bit significands): hh + hl + lh
//                with v_mfma_f32_16x16x32_f16, operand error ~2^-22.  FP16's
//                range is handled with exact power-of-two scales: each weight
//                column is scaled on upload so its largest entry lies in
//                [2^13, 2^14) (the low terms stay normal); the value and
//                Laplacian rows of each unit are scanned before the K loop and
//                scaled the same way; Jacobian rows take the fixed scale 2^6
//                (entries below 2^10 in magnitude stay finite, and their
//                absolute error floor is 2^-31).  The accumulators are unscaled
//                before the epilogue.  Value-only launches (no tangent rows)
//                use the BF16X6 kernel.
// Activations are split into 16-bit terms when staged into LDS; weights are
// constant and arrive pre-split as NS planes (GemmArgs::Wt_lo).
//
// K-to-lane mapping: within a 32-wide K chunk, lane group g = lane/16 owns
// k = 8g .. 8g+7.  For the BF16 16x16x32 MFMA this is one 16-byte LDS read per
// operand term; for the FP32 16x16x4 MFMA, step kk uses k = 8g + kk, so each
// lane reads its 8 operands with two 16-byte LDS reads.  LDS row stride is 36
// floats / 40 bf16, which keeps 16-byte reads of 8 consecutive rows on
// distinct banks.

#pragma once

#include <type_traits>

#include "kernels.h"

// 0: default tile table; else UPT*10 + NBW (benchmark override)
extern int g_gemm_variant;
// 4: no eight-wave tiles (benchmark override); 0: default
extern int g_gemm_nw;

namespace {

template <int PREC> struct NSplit { static constexpr int v = 1; };
template <> struct NSplit<PREC_BF16X3> { static constexpr int v = 2; };
template <> struct NSplit<PREC_BF16X6> { static constexpr int v = 3; };
template <> struct NSplit<PREC_FP16X3> { static constexpr int v = 2; };

template <typename T>
__device__ inline void split3(float x, T* t, int ns) {
  float r = x;
  for (int i = 0; i < ns; ++i) {
    const T h = (T)r;
    t[i] = h;
    r = r - (float)h;
  }
}

// Fixed power-of-two scale of Jacobian rows in PREC_FP16X3 (see above).
constexpr float kF16JScale = 64.f;

// MBU: 16-row blocks per unit (Laplacian mode); UPT: units per tile; NBW:
// 16-column blocks per wave.  BM = 16*MBU*UPT, BN = 16*NBW*NW (NW waves).
// Waves per SIMD the register allocation must leave room for: three for tiles
// of at most 12 accumulator blocks, two otherwise.
template <int MB, int NBW>
constexpr int gemm_min_waves() {
  return MB * NBW <= 12 ? 3 : 2;
}

// LNORM: the fused-LayerNorm variant (GemmArgs::Yln set; no head-split,
// broadcast or padding outputs), a kernel of its own so that the plain
// variant keeps its register allocation.
// NW: waves per workgroup (4, or 8 for a 256-column tile of two-block waves,
// whose operand tile all eight share).
template <int MBU, int UPT, int NBW, int PREC, bool LAP, bool LNORM = false, int NW = 4>
__global__ __launch_bounds__(64 * NW, (gemm_min_waves<MBU * UPT, NBW>())) void gemm_kernel(const GemmArgs a) {
  static_assert(NW == 4 || (NW == 8 && !LNORM), "four waves, or eight without the fused LayerNorm");
  constexpr int NT = 64 * NW;  // threads
  constexpr int MB = MBU * UPT;
  constexpr int BM = 16 * MB;
  constexpr int BN = 16 * NBW * NW;
  constexpr int KC = 32;
  constexpr bool F32 = PREC == PREC_F32;
  constexpr bool F16 = PREC == PREC_FP16X3;
  constexpr int NS = NSplit<PREC>::v;
  constexpr int LS = F32 ? 36 : 40;  // LDS row stride (elements)
  using TL = typename std::conditional<F32, float, typename std::conditional<F16, f16, bf16>::type>::type;
  constexpr int NSL = F32 ? 1 : NS;
  // LDS holds the activation tile of one K chunk, which all the waves read;
  // the epilogue reuses it to stage output rows.  The weights are not staged:
  // every wave owns its own columns, so each lane loads its MFMA B fragments
  // (8 consecutive k of one column) straight from global memory, where the
  // weights stay L2-resident.
  constexpr int A_BYTES = NSL * BM * LS * (int)sizeof(TL);
  // Two operand tiles when they fit 48 KB (three workgroups per CU): chunk
  // it+1 is staged into one while chunk it is multiplied from the other, so
  // the K loop needs one barrier per chunk.  Larger tiles use one buffer and
  // two barriers.
  constexpr bool DB = 2 * A_BYTES <= (NW == 8 ? 96 : 48) * 1024;  // eight waves: one workgroup per CU
  constexpr int NBUF = DB ? 2 : 1;
  constexpr int STG_BYTES = NW * 16 * (NBW * 16 + 4) * 4;
  constexpr int ARENA = NBUF * A_BYTES > STG_BYTES ? NBUF * A_BYTES : STG_BYTES;
  __shared__ __attribute__((aligned(16))) char arena[ARENA];
  constexpr int PLANE = BM * LS;  // elements of one operand plane
  TL* const Abase = (TL*)arena;
  // PREC_FP16X3: power-of-two scale of each tile row (1 in the other modes)
  __shared__ float rsc[F16 ? BM : 1], rsi[F16 ? BM : 1];

  // register prefetch: A chunk as float4 (8 per 32-wide row); B fragments
  constexpr int AV = (BM * 8 + NT - 1) / NT;
  f32x4 pa[AV];
  // B fragments, in two sets used alternately by consecutive chunks (see the
  // K loop); set 1 is used only where EARLY_B holds.
  f32x4 bn32[2][NBW][2];  // FP32: k = 8g .. 8g+7 of column nb
  uint4 bnb[2][NS][NBW];   // 16-bit planes: the same 8 k as 8 halves

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wcol = wave * NBW * 16;
  int rt, cb;
  if (a.linear_tiles) {
    const int ncb = (a.ncols + BN - 1) / BN;
    rt = blockIdx.x / ncb;
    cb = blockIdx.x % ncb;
  } else {
    // XCD x = b % 8 takes the contiguous row tiles [x*tpx, (x+1)*tpx) (the
    // grid is xcd_grid's), its column blocks of one row tile back to back:
    // the column blocks share the tile's X panel in that XCD's L2, and the
    // units of one walker (with their broadcast term G) stay on one XCD.
    const int ncb = (a.ncols + BN - 1) / BN;
    const int tpx = (int)gridDim.x / (kXcds * ncb);
    const int x = (int)blockIdx.x % kXcds, slot = (int)blockIdx.x / kXcds;
    cb = slot % ncb;
    rt = x * tpx + slot / ncb;
  }
  if ((long)rt * BM >= a.rows) return;
  const long row0 = (long)rt * BM;
  const int col0 = cb * BN;
  const float* X = (const float*)a.X + (long)blockIdx.z * a.sxb;
  // element offset of row gr in X (ld) or Y (ldy): row groups of rgrp rows
  // at group strides sxg / syg when rgrp > 0 (plain-row launches only)
  auto roff = [&](long gr, long ld, long sg) __attribute__((always_inline)) {
    if constexpr (!LAP) {
      if (a.rgrp > 0) {
        const long g = gr / a.rgrp;
        return g * sg + (gr - g * a.rgrp) * ld;
      }
    }
    return gr * ld;
  };
  const float* W = (const float*)a.Wt;
  // FP16X3 weights: two FP16 planes after the three BF16 planes, then the
  // inverse column scales and the column scales (powers of two)
  const TL* Wb = F16 ? (const TL*)((const bf16*)a.Wt_lo + 3 * a.wplane) : (const TL*)a.Wt_lo;
  const float* csinv = F16 ? (const float*)(Wb + 2 * a.wplane) : nullptr;

  // Operand loads are unconditional: rows past the end read the last row and
  // columns past the end read column 0, so no load sits behind a branch (the
  // compiler then counts outstanding loads exactly).  Such rows and columns
  // only reach accumulators that are never stored: rows are independent,
  // Laplacian-mode units are whole, and output columns >= ncols are skipped.
  // the element offsets of the thread's AV staging rows (and column), fixed
  // for the tile
  long xo[AV];
#pragma unroll
  for (int v = 0; v < AV; ++v) {
    int idx = tid + v * NT;
    if (idx >= BM * 8) idx -= BM * 8;  // duplicate of an in-tile load (not staged)
    const int r = idx >> 3, q = idx & 7;
    const long gr = row0 + r < a.rows ? row0 + r : (long)a.rows - 1;
    xo[v] = roff(gr, a.ldx, a.sxg) + q * 4;
  }
  auto gload = [&](int k0) __attribute__((always_inline)) {
#pragma clang loop unroll(full)
    for (int v = 0; v < AV; ++v) pa[v] = *(const f32x4*)(X + xo[v] + k0);
  };
  auto bload = [&](auto set, int k0) __attribute__((always_inline)) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      const int gc = col0 + wcol + nb * 16 + lr;
      const long gcl = gc < a.ncols ? gc : 0;
      if constexpr (F32) {
        const float* p = W + gcl * a.ldw + k0 + 8 * lg;
        bn32[S][nb][0] = *(const f32x4*)p;
        bn32[S][nb][1] = *(const f32x4*)(p + 4);
      } else {
#pragma unroll
        for (int s = 0; s < NS; ++s) bnb[S][s][nb] = *(const uint4*)(Wb + s * a.wplane + gcl * a.ldw + k0 + 8 * lg);
      }
    }
  };
  // the 16-bit fragments of column blocks [nb0, nb0 + nbn) only
  auto bload_part = [&](auto set, int k0, int nb0, int nbn) __attribute__((always_inline)) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int nb = nb0; nb < nb0 + nbn; ++nb) {
      const int gc = col0 + wcol + nb * 16 + lr;
      const long gcl = gc < a.ncols ? gc : 0;
#pragma unroll
      for (int s = 0; s < NS; ++s) bnb[S][s][nb] = *(const uint4*)(Wb + s * a.wplane + gcl * a.ldw + k0 + 8 * lg);
    }
  };
  auto lstore = [&](int buf) __attribute__((always_inline)) {
    TL* const Ab = Abase + buf * NSL * PLANE;
#pragma clang loop unroll(full)
    for (int v = 0; v < AV; ++v) {
      const int idx = tid + v * NT;
      const int r = idx >> 3, q = idx & 7;
      if ((v + 1) * NT > BM * 8 && idx >= BM * 8) continue;  // only a last, partial pass
      if constexpr (F32) {
        *(f32x4*)(&Ab[r * LS + q * 4]) = pa[v];
      } else {
        f32x4 x = pa[v];
        if constexpr (F16) {
          const float sc = rsc[r];
          x.x *= sc;
          x.y *= sc;
          x.z *= sc;
          x.w *= sc;
        }
        TL t[4][NS];
        split3(x.x, t[0], NS);
        split3(x.y, t[1], NS);
        split3(x.z, t[2], NS);
        split3(x.w, t[3], NS);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          TL h[4] = {t[0][s], t[1][s], t[2][s], t[3][s]};
          *(uint2*)(&Ab[s * PLANE + r * LS + q * 4]) = *(uint2*)h;
        }
      }
    }
  };

  // When the residual is a slice of this GEMM's own input (X[row][n] for
  // output column n < ncols), the K loop starts after the residual columns
  // and ends on them, so the epilogue's residual reads follow the row tile's
  // most recent loads of those columns.  Every column block of a row tile
  // walks K in the same order (they run back to back on one XCD and share
  // the X panel's loads in its L2).
  const int nk = a.K / KC;
  const bool res_in_x = a.Res != nullptr && (const void*)a.Res == a.X && a.ldres == a.ldx && a.ncols <= a.K;
  const int kst = res_in_x ? ((a.ncols + KC - 1) / KC) % nk : 0;
  auto kof = [&](int it) __attribute__((always_inline)) {
    const int c = it + kst;
    return (c >= nk ? c - nk : c) * KC;
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  // The first chunk's operands are requested before the row-scale scan, so
  // their latency and the scan's overlap (not in tiles of 32 accumulator
  // blocks, whose registers are all taken by then).
  constexpr bool FIRST_EARLY = MB * NBW < 32;
  if constexpr (FIRST_EARLY) {
    gload(kof(0));
    bload(S0{}, kof(0));
  }

  // PREC_FP16X3 row scales: every row takes the fixed Jacobian scale, then
  // the value and Laplacian rows of every unit (row % period 0 or 1) are
  // scanned over K, one wave per row.  The candidate rows of the tile are
  // enumerated from one modulo per workgroup.
  float csi[NBW], csf[NBW];  // inverse and plain column scales of the lane's columns (FP16X3)
  if constexpr (F16) {
    const int per = LAP ? 16 * MBU : (a.lrp > 0 ? a.lrp : 1);
    for (int r = tid; r < BM; r += NT) {
      const bool in = row0 + r < a.rows;
      rsc[r] = in ? kF16JScale : 1.f;
      rsi[r] = in ? 1.f / kF16JScale : 1.f;
    }
    __syncthreads();
    const int off = (int)(row0 % per);
    const int cpr = per > 1 ? 2 : 1;  // scanned rows per period
    const int ncand = cpr * ((BM + off + per - 1) / per);
    for (int q = wave; q < ncand; q += NW) {
      const int r = (q / cpr) * per + (q % cpr) - off;
      const long gr = row0 + r;
      if (r < 0 || r >= BM || gr >= a.rows) continue;
      float mx = 0.f;
      for (int k = lane * 4; k < a.K; k += 256) {
        const float4 v = *(const float4*)(X + roff(gr, a.ldx, a.sxg) + k);
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
      }
      mx = wave_max(mx);
      float sc = 1.f, si = 1.f;
      if (mx > 0.f && mx <= 3.0e38f) {
        int e;
        (void)frexpf(mx, &e);
        int sh = 14 - e;
        sh = sh < -60 ? -60 : (sh > 60 ? 60 : sh);
        sc = ldexpf(1.f, sh);
        si = ldexpf(1.f, -sh);
      }
      if (lane == 0) {
        rsc[r] = sc;
        rsi[r] = si;
      }
    }
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      const int gc = col0 + wcol + nb * 16 + lr;
      csi[nb] = gc < a.ncols ? csinv[gc] : 1.f;
      csf[nb] = gc < a.ncols ? csinv[a.ncols + gc] : 1.f;
    }
    __syncthreads();
  }

  f32x4 acc[MB][NBW];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NBW; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // Laplacian mode: the accumulators start from the broadcast term
  // G[walker, row-in-unit, col], so the epilogue needs no extra loads.
  // acc[mb][nb][v] is tile row mb*16 + 4*lg + v, column wcol + nb*16 + lr.
  // Units are whole (rows % RU == 0), so validity is per unit; columns past
  // ncols read a clamped address and add nothing.
  if constexpr (LAP) {
    int ncl[NBW];
    bool cin[NBW];
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      const int n = col0 + wcol + nb * 16 + lr;
      cin[nb] = n < a.ncols;
      ncl[nb] = cin[nb] ? n : a.ncols - 1;
    }
    if (!LNORM && a.G) {
#pragma unroll
      for (int u = 0; u < UPT; ++u) {
        const long unit = row0 / (16 * MBU) + u;
        if ((unit + 1) * (16 * MBU) <= a.rows) {
          const long w = a.wmaj ? unit / (a.n_up + a.n_dn) : walker_of_unit((int)unit, a.Bc, a.n_up, a.n_dn);
          const float* gu = a.G + w * (16 * MBU) * a.ldg;
          // loaded straight into the accumulators, all at once (no
          // temporaries); columns >= ncols read column ncols - 1 and are
          // never stored.  FP16X3 scales them after the first chunk's loads
          // are issued (below).
#pragma unroll
          for (int m = 0; m < MBU; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const float* gr = gu + (long)(m * 16 + 4 * lg + v) * a.ldg;
#pragma unroll
              for (int nb = 0; nb < NBW; ++nb) acc[u * MBU + m][nb][v] = gr[ncl[nb]];
            }
        }
      }
    }
  }

  // EARLY_B: the next chunk's B fragments are requested into the other set
  // before this chunk's products, so no fragment is ever copied and no wait
  // covers loads issued after them.  Otherwise (three planes, or tiles of
  // four column blocks per wave in FP32, where two sets do not fit the
  // register budget) one set is used and the next chunk's fragments are
  // requested after their last product.  SPLITB (FP16X3, four column blocks
  // per wave, two operand tiles): a chunk's products run as two column
  // halves, each over every row block (the A fragments are read from LDS
  // once per half), so the next chunk's fragments of the first half are
  // requested half a chunk earlier.
  constexpr bool EARLY_B = F32 ? NBW <= 2 : NS * NBW <= 4;
  constexpr bool SPLITB = F16 && DB && !EARLY_B && NBW == 4;
  constexpr int NHS = SPLITB ? 2 : 1;  // column parts of a chunk's products
  // products of one chunk from operand tile `buf` with B fragment set `set`;
  // SPLITB requests the next chunk's (at knext) first-half fragments after
  // the first half's products
  auto mfma_chunk = [&](auto set, int buf, int knext) __attribute__((always_inline)) {
    constexpr int S = decltype(set)::value;
    const TL* const Ab = Abase + buf * NSL * PLANE;
    if constexpr (F32) {
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const f32x4 a0 = *(const f32x4*)&Ab[(mb * 16 + lr) * LS + lg * 8];
        const f32x4 a1 = *(const f32x4*)&Ab[(mb * 16 + lr) * LS + lg * 8 + 4];
#pragma unroll
        for (int kk = 0; kk < 8; ++kk)
#pragma unroll
          for (int nb = 0; nb < NBW; ++nb)
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(kk < 4 ? a0[kk & 3] : a1[kk & 3],
                                                               bn32[S][nb][kk >> 2][kk & 3], acc[mb][nb], 0, 0, 0);
      }
    } else {
      constexpr int NH = NHS;
      constexpr int NBH = NBW / NH;
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        if (h > 0) {
          // the second half reads its A fragments again (kept from the first
          // half they would take 64 registers)
          __asm__ volatile("" ::: "memory");
          bload_part(set, knext, (h - 1) * NBH, NBH);
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          bf16x8 at[NS];
#pragma unroll
          for (int s = 0; s < NS; ++s) at[s] = *(const bf16x8*)&Ab[s * PLANE + (mb * 16 + lr) * LS + lg * 8];
#pragma unroll
          for (int nb = h * NBH; nb < (h + 1) * NBH; ++nb) {
            if constexpr (F16) {
              // small terms first
              const f16x8 a0 = __builtin_bit_cast(f16x8, at[0]), a1 = __builtin_bit_cast(f16x8, at[1]);
              const f16x8 b0 = __builtin_bit_cast(f16x8, bnb[S][0][nb]), b1 = __builtin_bit_cast(f16x8, bnb[S][1][nb]);
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b0, acc[mb][nb], 0, 0, 0);
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b1, acc[mb][nb], 0, 0, 0);
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b0, acc[mb][nb], 0, 0, 0);
              continue;
            }
            bf16x8 bt[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) bt[s] = __builtin_bit_cast(bf16x8, bnb[S][s][nb]);
            // small terms first
            if constexpr (NS == 3) {
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[1], bt[1], acc[mb][nb], 0, 0, 0);
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[2], acc[mb][nb], 0, 0, 0);
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[2], bt[0], acc[mb][nb], 0, 0, 0);
            }
            if constexpr (NS >= 2) {
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[1], acc[mb][nb], 0, 0, 0);
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[1], bt[0], acc[mb][nb], 0, 0, 0);
            }
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at[0], bt[0], acc[mb][nb], 0, 0, 0);
          }
        }
      }
    }
  };

  if constexpr (!FIRST_EARLY) {
    gload(kof(0));
    bload(S0{}, kof(0));
  }
  if constexpr (LAP) {
    if (!LNORM && a.G) {
      // The broadcast loads and the first chunk's are in flight together;
      // wait for all of them here (vmcnt 0), so the K loop starts with no
      // load outstanding from before it and its waits stay exact.
      __builtin_amdgcn_s_waitcnt(0x0F70);
      if constexpr (F16) {  // the accumulators carry the row and column scales
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float gs = rsc[mb * 16 + 4 * lg + v];
#pragma unroll
            for (int nb = 0; nb < NBW; ++nb) acc[mb][nb][v] *= gs * csf[nb];
          }
      }
    }
  }
  if constexpr (DB) {
    // chunk it is multiplied from tile it&1 while chunk it+1 is staged into
    // the other tile (its loads were issued one chunk earlier); the barrier
    // at the end of a step publishes tile (it+1)&1 and retires every read of
    // tile it&1 before step it+1 overwrites it.
    lstore(0);
    gload(kof(nk > 1 ? 1 : 0));
    __syncthreads();
    // Every load of a step is unconditional (past the last chunk they repeat
    // it, and the last step's staging write goes to the idle tile), so the
    // compiler's load counters agree on every path through the loop and no
    // wait covers loads issued after the ones it needs.
    auto kc = [&](int c) __attribute__((always_inline)) { return kof(c < nk ? c : nk - 1); };
    auto step = [&](auto set, int it) __attribute__((always_inline)) {
      constexpr int S = decltype(set)::value;
      if constexpr (EARLY_B) bload(std::integral_constant<int, 1 - S>{}, kc(it + 1));
      mfma_chunk(set, it & 1, kc(it + 1));
      if constexpr (SPLITB) {
        bload_part(set, kc(it + 1), NBW - NBW / NHS, NBW / NHS);
      } else if constexpr (!EARLY_B) {
        bload(set, kc(it + 1));
      }
      lstore((it + 1) & 1);
      gload(kc(it + 2));
      __syncthreads();
    };
    if constexpr (EARLY_B) {
      int it = 0;
      for (; it + 1 < nk; it += 2) {
        step(S0{}, it);
        step(S1{}, it + 1);
      }
      if (it < nk) step(S0{}, it);
    } else {
      for (int it = 0; it < nk; ++it) step(S0{}, it);
    }
  } else {
    for (int it = 0; it < nk; ++it) {
      lstore(0);
      __syncthreads();
      if (it + 1 < nk) gload(kof(it + 1));
      mfma_chunk(S0{}, 0, 0);
      if (it + 1 < nk) bload(S0{}, kof(it + 1));
      __syncthreads();
    }
  }

  if constexpr (F16) {  // remove the power-of-two row and column scales
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float ir = rsi[mb * 16 + 4 * lg + v];
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) acc[mb][nb][v] *= ir * csi[nb];
      }
  }

  // ---- epilogue (registers) ----
  // acc[mb][nb][v] holds tile row mb*16 + 4*lg + v, column wcol + nb*16 + lr.
  // A wave owns every row of the tile for its columns, so in Laplacian mode
  // the RU = 16*MBU rows of unit u are its row blocks u*MBU .. u*MBU+MBU-1 and
  // the tanh rule's column sum over the Jacobian rows is an in-thread sum plus
  // two shuffles across the lane groups.  Stores are 64-byte row segments.
  constexpr int RU = LAP ? 16 * MBU : 1;
  float* Y = (float*)a.Y + (long)blockIdx.z * a.syb;
  // With the fused LayerNorm the residual goes into the accumulators first
  // (pass 4 normalises the sum), and the store below adds none.
  constexpr bool LNF = LNORM;
  static_assert(!LNORM || (LAP && UPT == 1), "the fused LayerNorm needs whole units per tile");
  constexpr bool lnf = LNF;
  const float* Res = lnf ? nullptr : (const float*)a.Res;

  int ncol[NBW];
  bool cok[NBW];
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) {
    ncol[nb] = col0 + wcol + nb * 16 + lr;
    cok[nb] = ncol[nb] < a.ncols;
  }
  // pass 1 (Laplacian mode): bias on the value rows
  if (LAP && a.bias) {
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      const float bv = (cok[nb] && lg == 0) ? a.bias[cok[nb] ? ncol[nb] : 0] : 0.f;
#pragma unroll
      for (int u = 0; u < UPT; ++u) acc[u * MBU][nb][ROW_VAL] += bv;
    }
  }
  // pass 2 (Laplacian mode): tanh rule per (unit, column):
  //   value <- t, Laplacian <- d1*lap - 2 t d1 sum_J J^2, J <- d1*J, d1 = 1 - t^2
  if (LAP && a.act == ACT_TANH) {
    {
#pragma unroll
      for (int u = 0; u < UPT; ++u)
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) {
          float sq = 0.f;
#pragma unroll
          for (int m = 0; m < MBU; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const float j = acc[u * MBU + m][nb][v];
              const bool jrow = !(m == 0 && lg == 0 && v < ROW_J0);
              sq += jrow ? j * j : 0.f;
            }
          sq += __shfl_xor(sq, 16);
          sq += __shfl_xor(sq, 32);
          const float y0 = __shfl(acc[u * MBU][nb][ROW_VAL], lr);
          const float y1 = __shfl(acc[u * MBU][nb][ROW_LAP], lr);
          const float t = tanhf(y0);
          const float d1 = 1.f - t * t;
          const float lap = d1 * y1 - 2.f * t * d1 * sq;
#pragma unroll
          for (int m = 0; m < MBU; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u * MBU + m][nb][v] *= d1;
          if (lg == 0) {
            acc[u * MBU][nb][ROW_VAL] = t;
            acc[u * MBU][nb][ROW_LAP] = lap;
          }
        }
    }
  }
  if constexpr (LNF) {
    if (lnf && a.Res) {
      const float* rs = (const float*)a.Res;
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float* rp = rs + (row0 + mb * 16 + 4 * lg + v) * a.ldres;
#pragma unroll
          for (int nb = 0; nb < NBW; ++nb) acc[mb][nb][v] = (acc[mb][nb][v] + rp[ncol[nb]]) * a.res_scale;
        }
      }
    }
  }
  // pass 3: residual and store (row-major, or head-split for Psiformer Q/K/V).
  // Each wave stages one 16-row block of its columns at a time in its own
  // slice of the (now free) operand arena and writes it back with 16-byte
  // accesses (16 rows of 64*NBW contiguous bytes).  The block loop is not
  // unrolled, so the store code exists once; the staging write selects the
  // block's accumulators with compile-time indices.
  constexpr int WC = NBW * 16, SW = WC + 4, C4W = WC / 4, RPI = 64 / C4W;
  static_assert(NW * 16 * SW * 4 <= ARENA, "staging slices must fit the arena");
  static_assert(64 % C4W == 0 && 16 % RPI == 0, "store pattern");
  __syncthreads();  // every wave is done with the operand tiles
  float* stg = (float*)arena + wave * 16 * SW;
  const bool vec_io = (a.ldy % 4 == 0) && (!Res || a.ldres % 4 == 0) && (a.hs_H == 0 || a.hs_dh % 4 == 0);
  // The lane's 4-column group is the same for every row it stores.
  const int c4 = (lane % C4W) * 4, r_lane = lane / C4W;
  const int n = col0 + wcol + c4;
  const bool nok = n < a.ncols;
  const bool tail = LNORM ? false : !(n + 4 <= a.ncols && vec_io);
  // head-split destination: column part ((part*H + h)*N*RU)*dh + d is per lane,
  // the row part ((w*3*H*N + i)*RU + rr)*dh per row; Laplacian-mode units of the
  // tile are consecutive, so (w, i) of its first unit is computed once.
  // plain mode: the lane's four bias values, the same for every row it stores
  float bq[4] = {0.f, 0.f, 0.f, 0.f};
  if (!LAP && a.bias) {
#pragma unroll
    for (int q = 0; q < 4; ++q) bq[q] = n + q < a.ncols ? a.bias[n + q] : 0.f;
  }
  long hcol = 0;
  int hw0 = 0, hi0 = 0;
  if (!LNORM && a.hs_H > 0) {
    const int part = n / a.hs_D, hh = (n % a.hs_D) / a.hs_dh, dd = n % a.hs_dh;
    hcol = ((long)part * a.hs_H + hh) * a.hs_N * ((long)RU * a.hs_dh) + dd;
    if (LAP) {
      const int u0 = (int)(row0 / RU);
      hw0 = u0 / a.hs_N;
      hi0 = u0 % a.hs_N;
    }
  }
  // Residual rows (vectorised case) are requested one 16-row block ahead, so
  // a block's stores do not wait for its own residual loads.
  constexpr int IPB = 16 / RPI;  // rows per lane and block
  const bool res_vec = Res != nullptr && !tail;
  float4 rqn[IPB];
  auto load_res = [&](int mb, float4* out) __attribute__((always_inline)) {
#pragma unroll
    for (int it = 0; it < IPB; ++it) {
      const long grow = row0 + mb * 16 + r_lane + RPI * it;
      out[it] = (res_vec && grow < a.rows) ? *(const float4*)(Res + grow * a.ldres + n) : float4{0.f, 0.f, 0.f, 0.f};
    }
  };
  load_res(0, rqn);
#pragma unroll 1
  for (int mb0 = 0; mb0 < MB; ++mb0) {
    float4 rq[IPB];
#pragma unroll
    for (int it = 0; it < IPB; ++it) rq[it] = rqn[it];
    if (mb0 + 1 < MB) load_res(mb0 + 1, rqn);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
      if (mb == mb0) {
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb)
#pragma unroll
          for (int v = 0; v < 4; ++v) stg[(4 * lg + v) * SW + nb * 16 + lr] = acc[mb][nb][v];
      }
    wave_sync();
#pragma unroll
    for (int it = 0; it < 16 / RPI; ++it) {
      const int r = r_lane + RPI * it;
      const long grow = row0 + mb0 * 16 + r;
      if (grow >= a.rows) continue;
      if (!LNORM && !nok) {  // padding columns [ncols, zpad) are written as zeros
        if (n < a.zpad && a.hs_H == 0) {
          float* dz = Y + roff(grow, a.ldy, a.syg) + n;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (n + q < a.zpad) dz[q] = 0.f;
        }
        continue;
      }
      float y[4];
      *(float4*)y = *(const float4*)&stg[r * SW + c4];
      if (!LAP) {
        // plain mode: broadcast term, bias and activation
        if (a.G) {
          const long w = a.wmaj ? grow / (a.n_up + a.n_dn) : walker_of_unit((int)grow, a.Bc, a.n_up, a.n_dn);
          const float* gp = a.G + w * a.ldg + n;
          if (!tail && a.ldg % 4 == 0) {
            const float4 g4 = *(const float4*)gp;
            y[0] += g4.x;
            y[1] += g4.y;
            y[2] += g4.z;
            y[3] += g4.w;
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (n + q < a.ncols) y[q] += gp[q];
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          y[q] += bq[q];
          if (a.act == ACT_TANH) y[q] = tanhf(y[q]);
        }
      }
      float* dst;
      if (!LNORM && a.hs_H > 0) {
        int w, i, rr;
        if (LAP) {
          const int u = (mb0 * 16 + r) / RU;  // unit of the tile
          rr = (mb0 * 16 + r) % RU;
          i = hi0 + u;
          w = hw0;
          while (i >= a.hs_N) {
            i -= a.hs_N;
            ++w;
          }
        } else {
          w = (int)(grow / a.hs_N);
          i = (int)(grow % a.hs_N);
          rr = 0;
        }
        dst = Y + ((((long)w * 3 * a.hs_H) * a.hs_N + i) * RU + rr) * (long)a.hs_dh + hcol;
      } else {
        dst = Y + roff(grow, a.ldy, a.syg) + n;
        if (Res) {
          const float* rp = Res + grow * a.ldres + n;
          if (!tail) {
            const float4 r4 = rq[it];
            y[0] = (y[0] + r4.x) * a.res_scale;
            y[1] = (y[1] + r4.y) * a.res_scale;
            y[2] = (y[2] + r4.z) * a.res_scale;
            y[3] = (y[3] + r4.w) * a.res_scale;
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (n + q < a.ncols) y[q] = (y[q] + rp[q]) * a.res_scale;
          }
        }
      }
      if (!tail) {
        *(float4*)dst = *(float4*)y;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (n + q < a.ncols)
            dst[q] = y[q];
          else if (n + q < a.zpad && a.hs_H == 0)
            dst[q] = 0.f;
        }
      }
    }
    wave_sync();
  }

  // pass 4 (fused LayerNorm): Yln = LN(output) by the forward-Laplacian rule
  // of layernorm_lap_kernel.  The tile is one unit with every column; each
  // wave holds 64*NBW/4 columns of all its rows, so a row statistic is a sum
  // over the 16 lanes of a lane group (one DPP row), combined across the four
  // waves in LDS by one thread per row; the Laplacian row's sum over the
  // Jacobian rows is an in-thread sum plus two shuffles across the lane
  // groups.  Unit row r is acc[r / 16][.][r % 4] of lane group (r / 4) % 4.
  if constexpr (LNF) {
    if (lnf) {
      constexpr float invD = 1.f / (64 * NBW);
      float* const part = (float*)arena;  // [4][RU] per-wave row partials
      float* const rowv = part + 4 * RU;  // [RU] per-row result (mean, then js)
      float* const scal = rowv + RU;      // [8] unit partials and scalars
      static_assert((5 * RU + 8) * 4 <= ARENA, "LayerNorm reductions must fit the arena");
      auto gsum = [&](float t) __attribute__((always_inline)) {
        t += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(t), 0xB1, 0xF, 0xF, false));
        t += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(t), 0x4E, 0xF, 0xF, false));
        t += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(t), 0x141, 0xF, 0xF, false));
        return t + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(t), 0x140, 0xF, 0xF, false));
      };
      __syncthreads();  // every wave's staging slice is free
      // row means; the accumulators become the centred rows d_r (the output
      // itself was stored above)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float t = 0.f;
#pragma unroll
          for (int nb = 0; nb < NBW; ++nb) t += acc[mb][nb][v];
          t = gsum(t);
          if (lr == 0) part[wave * RU + mb * 16 + 4 * lg + v] = t;
        }
      __syncthreads();
      if (tid < RU) rowv[tid] = ((part[tid] + part[RU + tid]) + (part[2 * RU + tid] + part[3 * RU + tid])) * invD;
      __syncthreads();
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float m = rowv[mb * 16 + 4 * lg + v];
#pragma unroll
          for (int nb = 0; nb < NBW; ++nb) acc[mb][nb][v] -= m;
        }
      // centred value row of the lane's columns (row 0 sits in lane group 0)
      float d0[NBW];
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb) d0[nb] = __shfl(acc[0][nb][0], lr);
      // <d_r, d_0> per row; sum of |d_r|^2 over the Jacobian rows
      float sq = 0.f;
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = mb * 16 + 4 * lg + v;
          float t = 0.f, q = 0.f;
#pragma unroll
          for (int nb = 0; nb < NBW; ++nb) {
            t += acc[mb][nb][v] * d0[nb];
            q += acc[mb][nb][v] * acc[mb][nb][v];
          }
          sq += (r >= ROW_J0 && r < a.ln_R) ? q : 0.f;
          t = gsum(t);
          if (lr == 0) part[wave * RU + r] = t;
        }
      sq = wave_sum(sq);
      if (lane == 0) scal[wave] = sq;
      __syncthreads();
      // unit scalars: variance v, its Laplacian, s = (v + eps)^-1/2 and its
      // first two derivatives; per Jacobian row js_r = s' J_v,r with
      // J_v,r = 2/D <d_r, d_0> (one thread per row), and sum_r J_v,r^2
      auto dotr = [&](int r) __attribute__((always_inline)) {
        return (part[r] + part[RU + r]) + (part[2 * RU + r] + part[3 * RU + r]);
      };
      const float sqt = (scal[0] + scal[1]) + (scal[2] + scal[3]);
      const float lapv = 2.f * invD * (dotr(ROW_LAP) + sqt);
      const float ve = dotr(ROW_VAL) * invD + a.ln_eps;
      const float sv = rsqrtf(ve);
      const float s1 = -0.5f * sv / ve;
      const float s2 = 0.75f * sv / (ve * ve);
      float jv2 = 0.f;
      if (tid < RU) {
        const float jv = (tid >= ROW_J0 && tid < a.ln_R) ? 2.f * invD * dotr(tid) : 0.f;
        rowv[tid] = s1 * jv;
        jv2 = jv * jv;
      }
      jv2 = wave_sum(jv2);
      __syncthreads();  // every wave has read part[] and scal[0..3]
      if (lane == 0) scal[4 + wave] = jv2;
      __syncthreads();
      const float laps = s1 * lapv + s2 * ((scal[4] + scal[5]) + (scal[6] + scal[7]));
      float g[NBW], bb[NBW], cross[NBW];
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb) {
        g[nb] = a.ln_s[ncol[nb]];
        bb[nb] = a.ln_b[ncol[nb]];
        cross[nb] = 0.f;
      }
      // value and Jacobian rows; the Laplacian row waits for
      // cross = sum_r d_r js_r over the Jacobian rows
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = mb * 16 + 4 * lg + v;
          const float js = rowv[r];  // 0 outside the Jacobian rows
          const bool jr = r >= ROW_J0 && r < a.ln_R;
#pragma unroll
          for (int nb = 0; nb < NBW; ++nb) {
            const float d = acc[mb][nb][v];
            cross[nb] += d * js;
            if (r == ROW_VAL)
              acc[mb][nb][v] = g[nb] * (d * sv) + bb[nb];
            else if (r != ROW_LAP)
              acc[mb][nb][v] = jr ? g[nb] * (d * sv + d0[nb] * js) : 0.f;
          }
        }
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb) {
        float t = cross[nb] + __shfl_xor(cross[nb], 16);
        t += __shfl_xor(t, 32);
        if (lg == 0) acc[0][nb][ROW_LAP] = g[nb] * (acc[0][nb][ROW_LAP] * sv + d0[nb] * laps + 2.f * t);
      }
      __syncthreads();  // the reductions are read before the staging slices reuse the arena
#pragma unroll 1
      for (int mb0 = 0; mb0 < MB; ++mb0) {
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          if (mb == mb0) {
#pragma unroll
            for (int nb = 0; nb < NBW; ++nb)
#pragma unroll
              for (int v = 0; v < 4; ++v) stg[(4 * lg + v) * SW + nb * 16 + lr] = acc[mb][nb][v];
          }
        wave_sync();
#pragma unroll
        for (int it = 0; it < 16 / RPI; ++it) {
          const int r = r_lane + RPI * it;
          *(float4*)(a.Yln + (row0 + mb0 * 16 + r) * a.ncols + n) = *(const float4*)&stg[r * SW + c4];
        }
        wave_sync();
      }
    }
  }
}

template <int MBU, int UPT, int NBW, int PREC, bool LAP, bool LNORM = false, int NW = 4>
void launch_one(const GemmArgs& a, hipStream_t s) {
  constexpr int BM = 16 * MBU * UPT, BN = 16 * NBW * NW;
  const int ntiles = a.linear_tiles ? ceil_div(a.rows, BM) * ceil_div(a.ncols, BN)
                                    : xcd_grid(ceil_div(a.rows, BM), ceil_div(a.ncols, BN));
  dim3 grid(ntiles, 1, a.nbatch > 1 ? a.nbatch : 1);
  hipLaunchKernelGGL((gemm_kernel<MBU, UPT, NBW, PREC, LAP, LNORM, NW>), grid, dim3(64 * NW), 0, s, a);
}


// Tile shapes that exist per unit height; FP32 has no four-block (256-column)
// tiles: with the FP32 weight fragments they exceed 256 VGPRs and spill.
template <int MBU, int PREC>
bool launch_lap(const GemmArgs& a, int upt, int nbw, hipStream_t s) {
#define JQ_TRY(U, NB)                                                           \
  if constexpr (MBU * U <= 12 && MBU * U * NB <= 36 && !(PREC == PREC_F32 && NB == 4)) { \
    if (upt == U && nbw == NB) {                                                \
      launch_one<MBU, U, NB, PREC, true>(a, s);                                 \
      return true;                                                              \
    }                                                                           \
  }
  // the fused-LayerNorm launches: FP16X3, one 128-row unit and 256 columns per tile
  if constexpr (PREC == PREC_FP16X3 && MBU == 8) {
    if (a.Yln) {
      if (upt != 1 || nbw != 4) return false;
      launch_one<MBU, 1, 4, PREC, true, true>(a, s);
      return true;
    }
  }
  if (a.Yln) return false;
  // FP16X3 192-row tiles of two-block waves wider than 128 columns: eight
  // waves share one operand tile over 256 columns (the two 128-column blocks
  // of four waves each streamed the activation panel separately)
  if constexpr (PREC == PREC_FP16X3 && (MBU == 3 || MBU == 4 || MBU == 6 || MBU == 12)) {
    constexpr int U8 = 12 / MBU;
    if (upt == U8 && nbw == 2 && a.ncols > 128 && g_gemm_nw != 4) {
      launch_one<MBU, U8, 2, PREC, true, false, 8>(a, s);
      return true;
    }
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
// MI355X (K = 256-320, 256 output columns, jq_bench_gemm).  At 48-row units
// the 16-bit modes take four units per tile (the weight fragments of a chunk
// are reused over 12 row blocks), FP32 two.  FP16X3 at 128-row units takes
// 256-column tiles when the layer is wider than 128 columns (the operand tile
// is staged and split once for twice the columns).  FP16X3 layers wider than
// 128 columns at 48-, 64-, 96- and 192-row units run 192-row tiles of eight
// waves (R = 64: 2.41 -> 2.33 ms, R = 96: 2.78 -> 2.31 ms at K = 320, 256
// columns).
void default_cfg(int R, int prec, bool wide, int* upt, int* nbw) {
  *nbw = 2;
  switch (R) {
    case 16: *upt = 4; break;
    case 32: *upt = 2; break;
    case 48: *upt = prec == PREC_F32 ? 2 : 4; break;
    case 64: *upt = prec == PREC_FP16X3 && wide ? 3 : 2; break;  // FP16X3: 192-row eight-wave tiles
    case 96: *upt = prec == PREC_FP16X3 && wide ? 2 : 1; break;
    case 128:
      *upt = 1;
      if (prec == PREC_FP16X3 && wide) *nbw = 4;
      break;
    default: *upt = 1; break;
  }
}

template <int PREC>
int dispatch(const GemmArgs& a, hipStream_t s) {
  const bool wide = a.ncols > 128;
  if (a.R <= 1) {
    if (a.Yln) return -1;
    // two 128-column blocks share their X panel through the XCD's L2, which
    // beats one 256-column block at a third of the occupancy.  Launches of
    // fewer than two rounds of 128-row tiles over the CUs (the Metropolis
    // forward pass, the per-walker broadcast GEMMs there) take 64-row tiles,
    // so the last round is not a mostly idle tail.
    const bool wide_tile = g_gemm_variant > 0 ? (g_gemm_variant % 10) == 4 : false;
    const long tiles128 = (long)ceil_div(a.rows, 128) * ceil_div(a.ncols, 128) * (a.nbatch > 1 ? a.nbatch : 1);
    if constexpr (PREC != PREC_F32) {  // no 256-column FP32 tiles (they spill)
      if (wide_tile) {
        launch_one<8, 1, 4, PREC, false>(a, s);
        return 0;
      }
    }
    if (tiles128 < 2L * 3 * 256)
      launch_one<4, 1, 2, PREC, false>(a, s);
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
  if (a.Yln && !(upt == 1 && 64 * nbw == a.ncols)) return -1;  // LayerNorm needs whole rows per workgroup
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
namespace {

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

}  // namespace
