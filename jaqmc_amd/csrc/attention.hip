// Fused forward-Laplacian multi-head attention (Laplacian mode), one
// workgroup per (walker, head).
//
// The product rule of laplacian/primitives/core.py:494-500 applied to
// s = q.k, a = softmax_j(s), o = a.v (backbone/psiformer.py:29-60; q carries
// the 1/sqrt(dh) scale in its weights).  Q, K, V arrive head-split,
// [w][part][h][i][Rp][dh]: row 0 value, row 1 Laplacian, rows 2 .. R-1 the
// Jacobian over the 3N electron coordinates, zero rows up to Rp.  For query
// i, key j and tangent row c:
//   J_c[i][j] = Q_c[i].K_0[j] + Q_0[i].K_c[j]           (c >= 1; c = 1 is the
//                                                         first-order part of lap s)
//   lap s     = J_1 + 2 S3,  S3[i][j] = sum_{c>=2} Q_c[i].K_c[j]
//   mu_c[i]   = sum_j a J_c,  delta_c = J_c - mu_c,  Ja_c = a delta_c
//   lap a     = a (rj - nu + lap s - lam),  rj = sum_c delta_c^2,
//               nu = sum_j a rj,  lam = sum_j a lap s
//   O_0 = a V_0,  O_c = Ja_c V_0 + a V_c  (c >= 2),
//   O_1 = lap_a V_0 + a V_1 + 2 sum_{c>=2} Ja_c V_c.
// The workgroup stages the value rows Q_0, K_0, V_0 and the attention
// weights once, then streams the tangent rows through two LDS row buffers:
// row c is multiplied while row c + 1 is staged and row c + 2 is in flight
// from HBM.  For each c it forms the 16x16 tiles of J_c (and accumulates S3)
// with v_mfma_f32_16x16x4f32, centres and weights them in the MFMA layout
// (16-lane DPP row sums), forms O_c (and accumulates sum_c Ja_c V_c) and
// writes O_c straight to the attention output; two workgroup barriers per
// row.  The score tensors never leave LDS.
//
// The workgroup's waves (att_threads) share each product's 16x16 output tiles.  MFMA 16x16x4 (f32):
// lane (r = lane%16, g = lane/16) supplies A[r][k] and B[k][r]; result
// element v is row 4g + v, column r.  Lane group g owns the contiguous K range
// [g*K/4, (g+1)*K/4), so one 16-byte LDS read per operand feeds four MFMA
// steps (V is staged transposed, [d][j], to make its K index j contiguous).
// The tiles of a wave run as interleaved, independent MFMA chains.
//
// H3 (the FP16X3 arithmetic mode): Q and K rows are held in LDS as two FP16
// planes, x s = h + l with h = fp16(x s), l = fp16(x s - h), where s is a
// power of two per row that puts the row's largest |entry| in [2^13, 2^14)
// (the row's 16 staging lanes agree on it by an xor-shuffle maximum; rows that
// are all zero take s = 1).  The q.k products then run on
// v_mfma_f32_16x16x32_f16 as hh + hl + lh (operand error 2^-22), one
// accumulator per product, each multiplied by the inverse row scales before
// they are combined.  The FP16 planes take the LDS the FP32 rows took; V, the
// weights and every elementwise step stay FP32.

#include <type_traits>

#include "kernels.h"

namespace {

// Threads per workgroup: 16 waves (four per SIMD, to hide the LDS and
// barrier latency of the row loop) where the registers fit 128 per lane, 12
// waves for the FP16X3 variant (its three accumulators and FP16 fragment pairs
// need about 150).
constexpr int att_threads(bool h3) { return h3 ? 768 : 1024; }

template <int NT, int DH, bool H3>
struct AttnLayout {
  static constexpr int NP = 16 * NT;   // padded electrons
  static constexpr int SQ = DH + 4;    // FP32 Q, K rows [i][d] (floats)
  static constexpr int SQH = DH + 8;   // H3: FP16 Q, K rows (halves; 16-byte reads of 8 rows on distinct banks)
  static constexpr int QKF = H3 ? NP * SQH : NP * SQ;  // floats of one Q or K array (H3: both planes)
  static constexpr int SJ = NP + 4;    // [i][j] arrays and transposed V [d][j]
  // Q_0, K_0, two row buffers of (Q_c, K_c); V_0 and two row buffers of V_c;
  // a, Ja_c, J_1; the partial row sums; H3: inverse row scales of the six
  // Q / K arrays
  static constexpr int FLOATS = 6 * QKF + 3 * DH * SJ + 3 * NP * SJ + NT * NP + (H3 ? 6 * NP : 0);
};

template <int CTL>
__device__ inline float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTL, 0xF, 0xF, false));
}

// Maximum over the N consecutive lanes of an aligned lane group (N = 8, 16),
// within the 16-lane DPP row: quad_perm [1,0,3,2] and [2,3,0,1] give each
// lane its quad's maximum, row_half_mirror (lane i reads lane 7 - i of its
// half-row) joins the two quads of a half-row, row_mirror the two halves.
template <int N>
__device__ inline float group_max(float v) {
  static_assert(N == 8 || N == 16, "lane groups of 8 or 16");
  v = fmaxf(v, dpp_f<0xB1>(v));
  v = fmaxf(v, dpp_f<0x4E>(v));
  v = fmaxf(v, dpp_f<0x141>(v));
  if constexpr (N == 16) v = fmaxf(v, dpp_f<0x140>(v));
  return v;
}

// Sum over the 16 lanes of each DPP row (lanes 16r .. 16r+15), result in every
// lane of the row: quad_perm [1,0,3,2] and [2,3,0,1] (xor 1, xor 2), then
// row_half_mirror and row_mirror fold the quads and the halves.
__device__ inline float row16_sum(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x141, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x140, 0xF, 0xF, false));
  return v;
}

template <int NT, int DH, bool H3>
__global__ __launch_bounds__(att_threads(H3)) void attn_lap_kernel(AttnArgs a) {
  constexpr int ATT_THREADS = att_threads(H3);
  constexpr int ATT_WAVES = ATT_THREADS / 64;
  using L = AttnLayout<NT, DH, H3>;
  constexpr int NP = L::NP, SQ = L::SQ, SJ = L::SJ, SQH = L::SQH, QKF = L::QKF;
  constexpr int DT = DH / 16;                               // output column tiles
  constexpr int JT = NT * NT, OT = NT * DT;                  // score / output tiles
  constexpr int JTW = (JT + ATT_WAVES - 1) / ATT_WAVES, OTW = (OT + ATT_WAVES - 1) / ATT_WAVES;
  constexpr int RPW = NP / ATT_WAVES > 0 ? (NP + ATT_WAVES - 1) / ATT_WAVES : 1;  // query rows per wave
  constexpr int F4 = DH / 4;                                 // float4 per row
  constexpr int PER = (3 * NP * F4 + ATT_THREADS - 1) / ATT_THREADS;
  constexpr int KQ = DH / 4, KA = NP / 4;                    // K range per lane group (q.k / a.v)

  extern __shared__ float sm[];
  float* Q0 = sm;             // Q_0, then K_0 at + QKF
  float* K0 = Q0 + QKF;
  float* Qb = K0 + QKF;       // row c in buffer c % 2: Q_c at Qb + 2*QKF*(c%2), K_c after it
  float* V0 = Qb + 4 * QKF;   // V_0 transposed [d][j]
  float* Vb = V0 + DH * SJ;   // V_c transposed, buffer c % 2 at Vb + DH*SJ*(c%2)
  float* As = Vb + 2 * DH * SJ;  // attention weights a[i][j]
  float* Jt = As + NP * SJ;   // Ja_c (then S3, then lap a)
  float* J1 = Jt + NP * SJ;   // J_1
  float* Mp = J1 + NP * SJ;   // per column tile partial row sums sum_j a J_c  [Qt][i]
  float* RJ = Q0;             // after the tangent loop: rj [i][j]
  // H3: inverse row scales, one NP array per Q / K array, in the arrays' order
  // (Q_0, K_0, Q buffer 0, K buffer 0, Q buffer 1, K buffer 1)
  // (slot s: the Q / K array at Q0 + s * QKF; Q_0 0, K_0 1, buffer b 2 + 2b, 3 + 2b)
  float* IS = Mp + NT * NP;
  auto isc = [&](int slot) __attribute__((always_inline)) { return IS + slot * NP; };

  const int N = a.N, RP = a.Rp, R = a.R;
  const int z = blockIdx.x, w = z / a.H, h = z % a.H;
  const long blk = (long)N * RP * DH;
  const float* qsrc = a.QKV + ((long)(w * 3) * a.H + h) * blk;  // K and V follow at + H*blk, + 2*H*blk
  const long pstride = (long)a.H * blk;
  float* O = a.O + (long)w * N * RP * a.ldo + (long)h * DH;  // row (i, c) at ((i*RP + c) * ldo)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lg = lane >> 4;

  for (int e = tid; e < L::FLOATS; e += ATT_THREADS) sm[e] = 0.f;  // padding rows and columns stay zero

  // row c of Q, K, V (rows i < N) through registers.  Item `it` is float4 q
  // of row i of part m.  Q and K items run along the row (16-byte accesses to
  // consecutive addresses); V items run four along the row, then down the
  // rows, so that the transposed LDS writes of a wave hit 16 consecutive
  // columns j per quarter-row instead of two banks.
  static_assert(F4 % 4 == 0, "head dimension must be a multiple of 16");
  auto item = [&](int it, int& m, int& i, int& q) __attribute__((always_inline)) {
    m = it / (NP * F4);
    const int rem = it % (NP * F4);
    if (m < 2) {
      i = rem / F4;
      q = rem % F4;
    } else {
      const int rest = rem >> 2;
      i = rest % NP;
      q = (rest / NP) * 4 + (rem & 3);
    }
  };
  float4 pf[PER];
  auto fetch = [&](int c) __attribute__((always_inline)) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int m, i, q;
      item(tid + ATT_THREADS * p, m, i, q);
      pf[p] = (m < 3 && i < N && c < R) ? *(const float4*)(qsrc + m * pstride + ((long)i * RP + c) * DH + 4 * q)
                                        : float4{0.f, 0.f, 0.f, 0.f};
    }
  };
  // qs: the slot of a Q array (its K array follows); dv: a transposed V array
  auto commit = [&](int qs, float* dv) __attribute__((always_inline)) {
    float* const dq = Q0 + qs * QKF;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int m, i, q;
      item(tid + ATT_THREADS * p, m, i, q);
      if constexpr (H3) {
        // the F4 lanes of a Q / K row form an aligned lane group: they agree
        // on the row's power-of-two scale, then write their four entries as
        // FP16 high and low parts (every lane takes part in the shuffles)
        const float4 x = pf[p];
        float mx = (m < 2 && i < N) ? fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w))) : 0.f;
        mx = group_max<F4>(mx);
        if (m < 2 && i < N) {
          float sc = 1.f, si = 1.f;
          if (mx > 0.f && mx <= 3.0e38f) {
            int e;
            (void)frexpf(mx, &e);
            int sh = 14 - e;
            sh = sh < -60 ? -60 : (sh > 60 ? 60 : sh);
            sc = ldexpf(1.f, sh);
            si = ldexpf(1.f, -sh);
          }
          const float xs[4] = {x.x * sc, x.y * sc, x.z * sc, x.w * sc};
          f16 hi[4], lo[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            hi[u] = (f16)xs[u];
            lo[u] = (f16)(xs[u] - (float)hi[u]);
          }
          f16* base = (f16*)(dq + m * QKF);  // plane 0 rows, then plane 1 rows
          *(uint2*)(base + i * SQH + 4 * q) = *(const uint2*)hi;
          *(uint2*)(base + NP * SQH + i * SQH + 4 * q) = *(const uint2*)lo;
          if (q == 0) isc(qs + m)[i] = si;
        }
        if (m == 2 && i < N) {
          dv[(4 * q + 0) * SJ + i] = x.x;
          dv[(4 * q + 1) * SJ + i] = x.y;
          dv[(4 * q + 2) * SJ + i] = x.z;
          dv[(4 * q + 3) * SJ + i] = x.w;
        }
        continue;
      }
      if (m < 2 && i < N) {
        *(float4*)(dq + (m * NP + i) * SQ + 4 * q) = pf[p];
      } else if (m == 2 && i < N) {
        dv[(4 * q + 0) * SJ + i] = pf[p].x;
        dv[(4 * q + 1) * SJ + i] = pf[p].y;
        dv[(4 * q + 2) * SJ + i] = pf[p].z;
        dv[(4 * q + 3) * SJ + i] = pf[p].w;
      }
    }
  };
  // A row (16P + lr) / B row (16Q + lr) of an [row][k] array, four k from lane group lg's range
  auto rd4 = [&](const float* M, int ld, int row, int kbase, int g) __attribute__((always_inline)) {
    return *(const float4*)(M + row * ld + kbase + 4 * g);
  };
  auto mfma4 = [&](f32x4& acc, const float4& x, const float4& y) __attribute__((always_inline)) {
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, y.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, y.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.z, y.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w, y.w, acc, 0, 0, 0);
  };
  // H3: FP16 (high, low) fragments of row `row` of a Q / K array, k = 32 ks + 8 lg .. + 7
  struct Frag {
    f16x8 h, l;
  };
  auto rdh = [&](const float* arr, int row, int ks) __attribute__((always_inline)) {
    const f16* b = (const f16*)arr + row * SQH + 32 * ks + 8 * lg;
    return Frag{*(const f16x8*)b, *(const f16x8*)(b + NP * SQH)};
  };
  auto mf3 = [&](f32x4& acc, const Frag& x, const Frag& y) __attribute__((always_inline)) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(x.l, y.h, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(x.h, y.l, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(x.h, y.h, acc, 0, 0, 0);
  };
  auto put_tile = [&](float* D, const f32x4& acc, int P, int Qt) __attribute__((always_inline)) {
#pragma unroll
    for (int v = 0; v < 4; ++v) D[(16 * P + 4 * lg + v) * SJ + 16 * Qt + lr] = acc[v];
  };
  auto store_o = [&](const f32x4& acc, int P, int Dq, int c) __attribute__((always_inline)) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 16 * P + 4 * lg + v;
      if (i < N) O[((long)i * RP + c) * a.ldo + 16 * Dq + lr] = acc[v];
    }
  };
  // output tiles of a.v-type products: acc[t] += A[i][:] . Vt[d][:] for the wave's tiles
  auto av_tiles = [&](f32x4* acc, const float* A, const float* Vt) __attribute__((always_inline)) {
#pragma unroll
    for (int g = 0; g < KA / 4; ++g)
#pragma unroll
      for (int t = 0; t < OTW; ++t) {
        const int tile = wave + ATT_WAVES * t;
        if (tile < OT) {
          const int P = tile / DT, Dq = tile % DT;
          mfma4(acc[t], rd4(A, SJ, 16 * P + lr, lg * KA, g), rd4(Vt, SJ, 16 * Dq + lr, lg * KA, g));
        }
      }
  };
  const bool jok = lane < N;

  // ---- value rows and the attention weights ----
  fetch(0);
  __syncthreads();  // LDS zeroed
  commit(0, V0);
  fetch(1);
  __syncthreads();
  {
    f32x4 acc[JTW];
#pragma unroll
    for (int t = 0; t < JTW; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (H3) {
#pragma unroll
      for (int ks = 0; ks < DH / 32; ++ks)
#pragma unroll
        for (int t = 0; t < JTW; ++t) {
          const int tile = wave + ATT_WAVES * t;
          if (tile < JT) mf3(acc[t], rdh(Q0, 16 * (tile / NT) + lr, ks), rdh(K0, 16 * (tile % NT) + lr, ks));
        }
      const float* iq = isc(0);
      const float* ik = isc(1);
#pragma unroll
      for (int t = 0; t < JTW; ++t) {
        const int tile = wave + ATT_WAVES * t;
        if (tile < JT) {
          const float kj = ik[16 * (tile % NT) + lr];
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[t][v] *= iq[16 * (tile / NT) + 4 * lg + v] * kj;
        }
      }
    } else {
#pragma unroll
      for (int g = 0; g < KQ / 4; ++g)
#pragma unroll
        for (int t = 0; t < JTW; ++t) {
          const int tile = wave + ATT_WAVES * t;
          if (tile < JT)
            mfma4(acc[t], rd4(Q0, SQ, 16 * (tile / NT) + lr, lg * KQ, g), rd4(K0, SQ, 16 * (tile % NT) + lr, lg * KQ, g));
        }
    }
#pragma unroll
    for (int t = 0; t < JTW; ++t) {
      const int tile = wave + ATT_WAVES * t;
      if (tile < JT) put_tile(As, acc[t], tile / NT, tile % NT);
    }
  }
  commit(4, Vb + DH * SJ);  // row 1 into buffer 1
  fetch(2);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < RPW; ++t) {
    const int i = wave + ATT_WAVES * t;
    if (i < N) {
      const float s = jok ? As[i * SJ + lane] : -INFINITY;
      float m = s;
      for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
      const float e = jok ? expf(s - m) : 0.f;
      const float zs = wave_sum(e);
      if (lane < NP) As[i * SJ + lane] = e / zs;
    }
  }

  // ---- tangent rows ----
  // The score tiles stay in the MFMA layout through the centring: a wave
  // keeps its J tiles' a values, the rj = sum_c delta_c^2 accumulators and the
  // tiles' validity in registers; row sums sum_j a J_c are 16-lane DPP sums
  // per tile, combined over the NT column tiles through LDS.
  f32x4 s3[JTW], l6[OTW], rjt[JTW], at[JTW];
#pragma unroll
  for (int t = 0; t < JTW; ++t) {
    s3[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    rjt[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int t = 0; t < OTW; ++t) l6[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  __syncthreads();  // attention weights final
#pragma unroll
  for (int t = 0; t < JTW; ++t) {
    const int tile = wave + ATT_WAVES * t;
    if (tile < JT) {
      const int P = tile / NT, Qt = tile % NT;
#pragma unroll
      for (int v = 0; v < 4; ++v) at[t][v] = As[(16 * P + 4 * lg + v) * SJ + 16 * Qt + lr];
    }
  }

  // Row c lives in buffer c % 2.  Per row: J products -> partial row sums ->
  // barrier -> centring, commit of row c + 1 into the other buffer, fetch of
  // row c + 2 -> barrier -> O products.  The other buffer's last readers (the
  // products of row c - 1) are behind the first barrier, and row c + 1 is in
  // LDS before the second.
  for (int c = 1; c < R; ++c) {
    const int b = c & 1;
    const float* Qc = Qb + 2 * QKF * b;
    const float* Kc = Qc + QKF;
    const float* Vc = Vb + DH * SJ * b;
    // J_c = Q_c K_0^T + Q_0 K_c^T  and  S3 += Q_c K_c^T
    f32x4 acc[JTW];
#pragma unroll
    for (int t = 0; t < JTW; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bool cross = c >= ROW_J0;
    if constexpr (H3) {
      // one accumulator per product (each carries its own row scales)
      f32x4 a2[JTW], a3[JTW];
#pragma unroll
      for (int t = 0; t < JTW; ++t) a2[t] = a3[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < DH / 32; ++ks)
#pragma unroll
        for (int t = 0; t < JTW; ++t) {
          const int tile = wave + ATT_WAVES * t;
          if (tile < JT) {
            const int P = tile / NT, Qt = tile % NT;
            const Frag qc = rdh(Qc, 16 * P + lr, ks), q0 = rdh(Q0, 16 * P + lr, ks);
            const Frag k0 = rdh(K0, 16 * Qt + lr, ks), kc = rdh(Kc, 16 * Qt + lr, ks);
            mf3(acc[t], qc, k0);
            mf3(a2[t], q0, kc);
            if (cross) mf3(a3[t], qc, kc);
          }
        }
      const float *iqc = isc(2 + 2 * b), *ikc = isc(3 + 2 * b), *iq0 = isc(0), *ik0 = isc(1);
#pragma unroll
      for (int t = 0; t < JTW; ++t) {
        const int tile = wave + ATT_WAVES * t;
        if (tile < JT) {
          const int P = tile / NT, Qt = tile % NT;
          const int j = 16 * Qt + lr;
          const float sk0 = ik0[j], skc = ikc[j];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = 16 * P + 4 * lg + v;
            const float sqc = iqc[i];
            acc[t][v] = acc[t][v] * (sqc * sk0) + a2[t][v] * (iq0[i] * skc);
            if (cross) s3[t][v] += a3[t][v] * (sqc * skc);
          }
        }
      }
    } else {
#pragma unroll
      for (int g = 0; g < KQ / 4; ++g)
#pragma unroll
        for (int t = 0; t < JTW; ++t) {
          const int tile = wave + ATT_WAVES * t;
          if (tile < JT) {
            const int P = tile / NT, Qt = tile % NT;
            const float4 qc = rd4(Qc, SQ, 16 * P + lr, lg * KQ, g), q0 = rd4(Q0, SQ, 16 * P + lr, lg * KQ, g);
            const float4 k0 = rd4(K0, SQ, 16 * Qt + lr, lg * KQ, g), kc = rd4(Kc, SQ, 16 * Qt + lr, lg * KQ, g);
            mfma4(acc[t], qc, k0);
            mfma4(acc[t], q0, kc);
            if (cross) mfma4(s3[t], qc, kc);
          }
        }
    }
    if (c >= ROW_J0) {
#pragma unroll
      for (int t = 0; t < JTW; ++t) {
        const int tile = wave + ATT_WAVES * t;
        if (tile < JT) {
          const int P = tile / NT, Qt = tile % NT;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float rs = row16_sum(at[t][v] * acc[t][v]);
            if (lr == 0) Mp[Qt * NP + 16 * P + 4 * lg + v] = rs;
          }
        }
      }
    }
    __syncthreads();
    if (c == 1) {
      // J_1 is kept for the Laplacian row; V_1 enters it after the loop
#pragma unroll
      for (int t = 0; t < JTW; ++t) {
        const int tile = wave + ATT_WAVES * t;
        if (tile < JT) put_tile(J1, acc[t], tile / NT, tile % NT);
      }
    } else {
      // centre and weight: Ja_c = a (J_c - mu_c), rj += (J_c - mu_c)^2
#pragma unroll
      for (int t = 0; t < JTW; ++t) {
        const int tile = wave + ATT_WAVES * t;
        if (tile < JT) {
          const int P = tile / NT, Qt = tile % NT;
          const bool jin = 16 * Qt + lr < N;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = 16 * P + 4 * lg + v;
            float mu = 0.f;
#pragma unroll
            for (int q = 0; q < NT; ++q) mu += Mp[q * NP + i];
            const float d = (jin && i < N) ? acc[t][v] - mu : 0.f;
            rjt[t][v] += d * d;
            acc[t][v] = at[t][v] * d;
          }
          put_tile(Jt, acc[t], P, Qt);
        }
      }
    }
    if (c + 1 < R) {
      commit(2 + 2 * (b ^ 1), Vb + DH * SJ * (b ^ 1));
      fetch(c + 2);
    }
    __syncthreads();
    if (c == 1) continue;
    {
      // O_c = Ja_c V_0 + a V_c,  L6 += Ja_c V_c
      f32x4 acc[OTW];
#pragma unroll
      for (int t = 0; t < OTW; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int g = 0; g < KA / 4; ++g)
#pragma unroll
        for (int t = 0; t < OTW; ++t) {
          const int tile = wave + ATT_WAVES * t;
          if (tile < OT) {
            const int P = tile / DT, Dq = tile % DT;
            const float4 ja = rd4(Jt, SJ, 16 * P + lr, lg * KA, g), aa = rd4(As, SJ, 16 * P + lr, lg * KA, g);
            const float4 v0 = rd4(V0, SJ, 16 * Dq + lr, lg * KA, g), vc = rd4(Vc, SJ, 16 * Dq + lr, lg * KA, g);
            mfma4(acc[t], ja, v0);
            mfma4(acc[t], aa, vc);
            mfma4(l6[t], ja, vc);
          }
        }
#pragma unroll
      for (int t = 0; t < OTW; ++t) {
        const int tile = wave + ATT_WAVES * t;
        if (tile < OT) store_o(acc[t], tile / DT, tile % DT, c);
      }
    }
  }

  // ---- Laplacian row and value row of the output ----
  __syncthreads();
  fetch(1);  // V_1 (and Q_1, K_1, unused)
#pragma unroll
  for (int t = 0; t < JTW; ++t) {
    const int tile = wave + ATT_WAVES * t;
    if (tile < JT) {
      put_tile(Jt, s3[t], tile / NT, tile % NT);
      put_tile(RJ, rjt[t], tile / NT, tile % NT);
    }
  }
  commit(2, Vb);  // row 1 into buffer 0 (its readers are done)
  __syncthreads();
  // lap s = J_1 + 2 S3; lap a = a (rj - nu + lap s - lam)
#pragma unroll
  for (int t = 0; t < RPW; ++t) {
    const int i = wave + ATT_WAVES * t;
    if (i < N) {
      const float av = jok ? As[i * SJ + lane] : 0.f;
      const float ls = jok ? J1[i * SJ + lane] + 2.f * Jt[i * SJ + lane] : 0.f;
      const float r = jok ? RJ[i * SJ + lane] : 0.f;
      const float nu = wave_sum(av * r), lam = wave_sum(av * ls);
      if (jok) Jt[i * SJ + lane] = av * (r - nu + ls - lam);
    }
  }
  __syncthreads();
  {
    f32x4 o0[OTW], o1[OTW];
#pragma unroll
    for (int t = 0; t < OTW; ++t) {
      o0[t] = f32x4{0.f, 0.f, 0.f, 0.f};
      o1[t] = f32x4{2.f * l6[t][0], 2.f * l6[t][1], 2.f * l6[t][2], 2.f * l6[t][3]};
    }
    av_tiles(o0, As, V0);
    av_tiles(o1, Jt, V0);
    av_tiles(o1, As, Vb);
#pragma unroll
    for (int t = 0; t < OTW; ++t) {
      const int tile = wave + ATT_WAVES * t;
      if (tile < OT) {
        store_o(o0[t], tile / DT, tile % DT, ROW_VAL);
        store_o(o1[t], tile / DT, tile % DT, ROW_LAP);
      }
    }
  }
  // padding rows R .. Rp-1 of the output are zero
  for (int e = tid; e < N * (RP - R) * DH; e += ATT_THREADS) {
    const int d = e % DH, rc = e / DH, i = rc / (RP - R), c = R + rc % (RP - R);
    O[((long)i * RP + c) * a.ldo + d] = 0.f;
  }
}

template <int NT, int DH, bool H3>
int launch_one(const AttnArgs& a, hipStream_t s) {
  const size_t smem = (size_t)AttnLayout<NT, DH, H3>::FLOATS * 4;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)attn_lap_kernel<NT, DH, H3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem);
    attr = true;
  }
  hipLaunchKernelGGL((attn_lap_kernel<NT, DH, H3>), dim3(a.nz), dim3(att_threads(H3)), smem, s, a);
  return 0;
}

template <bool H3>
int launch_mode(const AttnArgs& a, hipStream_t s) {
  const int nt = (a.N + 15) / 16;
  if (a.dh == 64) {
    if (nt == 1) return launch_one<1, 64, H3>(a, s);
    if (nt == 2) return launch_one<2, 64, H3>(a, s);
    return launch_one<3, 64, H3>(a, s);
  }
  if (nt == 1) return launch_one<1, 32, H3>(a, s);
  if (nt == 2) return launch_one<2, 32, H3>(a, s);
  return launch_one<3, 32, H3>(a, s);
}

}  // namespace

// ---------------------------------------------------------------------------
// Value-only attention (the Metropolis forward pass): s = q.k (q carries the
// 1/sqrt(dh) scale), a = softmax_j(s), o = a.v (backbone/psiformer.py:29-60),
// FP32 on the VALU, one workgroup per (walker, head), nothing but Q, K, V and
// O in HBM.  The head's Q, K, V rows ([part][h][i][dh] of the head-split
// projection, 16-byte loads) and the score matrix live in LDS (rows padded by
// 4 floats).  Scores: each thread forms a 2 x 4 block (two query rows, four
// key rows) from 16-byte LDS reads; softmax: one wavefront per query row
// (lane j holds key j, DPP maximum and sum); output: each thread forms a
// 2 x 4 block (two query rows, four features) over the keys.
namespace {

template <int DH>
__global__ __launch_bounds__(256) void attn_fwd_kernel(const float* __restrict__ QKV, float* __restrict__ O, int ldo,
                                                       int N, int H) {
  constexpr int SR = DH + 4;  // Q, K, V row stride (floats)
  constexpr int D4 = DH / 4;
  extern __shared__ float fsm[];
  const int NP = (N + 3) & ~3, SS = NP + 4;
  float* Qs = fsm;
  float* Ks = Qs + N * SR;
  float* Vs = Ks + N * SR;
  float* Ss = Vs + N * SR;  // [i][j], stride SS
  const int z = blockIdx.x, w = z / H, h = z % H;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long blk = (long)N * DH;  // one (part, head) block
  const float* src = QKV + ((long)w * 3 * H + h) * blk;
  for (int e = tid; e < 3 * N * D4; e += 256) {
    const int part = e / (N * D4), r = e % (N * D4), i = r / D4, q = r % D4;
    const float4 v = *(const float4*)(src + (long)part * H * blk + (long)i * DH + 4 * q);
    *(float4*)&Qs[part * N * SR + i * SR + 4 * q] = v;
  }
  __syncthreads();
  // scores, 2 x 4 blocks; rows past N read row N - 1 and are not stored
  const int NI = (N + 1) / 2, NJ = NP / 4;
  for (int it = tid; it < NI * NJ; it += 256) {
    const int i0 = 2 * (it / NJ), j0 = 4 * (it % NJ);
    const int i1 = min(i0 + 1, N - 1);
    int jr[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) jr[q] = min(j0 + q, N - 1);
    float acc[2][4] = {};
#pragma unroll 4
    for (int d = 0; d < DH; d += 4) {
      const float4 a0 = *(const float4*)&Qs[i0 * SR + d], a1 = *(const float4*)&Qs[i1 * SR + d];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 b = *(const float4*)&Ks[jr[q] * SR + d];
        acc[0][q] += a0.x * b.x + a0.y * b.y + a0.z * b.z + a0.w * b.w;
        acc[1][q] += a1.x * b.x + a1.y * b.y + a1.z * b.z + a1.w * b.w;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      Ss[i0 * SS + j0 + q] = acc[0][q];
      if (i0 + 1 < N) Ss[(i0 + 1) * SS + j0 + q] = acc[1][q];
    }
  }
  __syncthreads();
  // softmax over the keys of each query row (N <= 64: lane j holds key j),
  // two rows at a time; maximum and sum are 16-lane DPP reductions combined
  // through scalar registers
  auto rl = [](float v, int l) __attribute__((always_inline)) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
  };
  auto dpp = [](float v, auto ctl) __attribute__((always_inline)) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), decltype(ctl)::value, 0xF,
                                                              0xF, false));
  };
  using Q1 = std::integral_constant<int, 0xB1>;
  using Q2 = std::integral_constant<int, 0x4E>;
  using HM = std::integral_constant<int, 0x141>;
  using RM = std::integral_constant<int, 0x140>;
  auto wmax = [&](float v) __attribute__((always_inline)) {
    v = fmaxf(v, dpp(v, Q1{}));
    v = fmaxf(v, dpp(v, Q2{}));
    v = fmaxf(v, dpp(v, HM{}));
    v = fmaxf(v, dpp(v, RM{}));
    return fmaxf(fmaxf(rl(v, 0), rl(v, 16)), fmaxf(rl(v, 32), rl(v, 48)));
  };
  auto wsum = [&](float v) __attribute__((always_inline)) {
    v = row16_sum(v);
    return (rl(v, 0) + rl(v, 16)) + (rl(v, 32) + rl(v, 48));
  };
  for (int i0 = 2 * wave; i0 < N; i0 += 8) {
    const int i1 = i0 + 1 < N ? i0 + 1 : i0;
    const float s0 = lane < N ? Ss[i0 * SS + lane] : -INFINITY;
    const float s1 = lane < N ? Ss[i1 * SS + lane] : -INFINITY;
    const float m0 = wmax(s0), m1 = wmax(s1);
    const float e0 = lane < N ? expf(s0 - m0) : 0.f, e1 = lane < N ? expf(s1 - m1) : 0.f;
    const float sum0 = wsum(e0), sum1 = wsum(e1);
    if (lane < NP) {
      Ss[i0 * SS + lane] = lane < N ? e0 / sum0 : 0.f;
      if (i0 + 1 < N) Ss[(i0 + 1) * SS + lane] = lane < N ? e1 / sum1 : 0.f;
    }
  }
  __syncthreads();
  // output rows, 2 x 4 blocks over the features
  float* ob = O + (long)w * N * ldo + h * DH;
  for (int it = tid; it < NI * D4; it += 256) {
    const int i0 = 2 * (it / D4), d = 4 * (it % D4);
    const int i1 = min(i0 + 1, N - 1);
    float4 o0 = {0.f, 0.f, 0.f, 0.f}, o1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int j = 0; j < N; ++j) {
      const float p0 = Ss[i0 * SS + j], p1 = Ss[i1 * SS + j];
      const float4 v = *(const float4*)&Vs[j * SR + d];
      o0.x += p0 * v.x;
      o0.y += p0 * v.y;
      o0.z += p0 * v.z;
      o0.w += p0 * v.w;
      o1.x += p1 * v.x;
      o1.y += p1 * v.y;
      o1.z += p1 * v.z;
      o1.w += p1 * v.w;
    }
    *(float4*)(ob + (long)i0 * ldo + d) = o0;
    if (i0 + 1 < N) *(float4*)(ob + (long)(i0 + 1) * ldo + d) = o1;
  }
}

}  // namespace

bool attn_lap_supported(int N, int dh) { return N >= 1 && N <= 48 && (dh == 32 || dh == 64); }

bool attn_fwd_supported(int N, int dh) { return N >= 1 && N <= 64 && (dh == 32 || dh == 64); }

int launch_attn_fwd(const AttnArgs& a, hipStream_t s) {
  if (!attn_fwd_supported(a.N, a.dh) || a.nz <= 0 || a.Rp != 1 || a.ldo % 4) return -1;
  const int NP = (a.N + 3) & ~3;
  const size_t smem = ((size_t)3 * a.N * (a.dh + 4) + (size_t)a.N * (NP + 4)) * sizeof(float);  // <= 70 KB
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)attn_fwd_kernel<32>, hipFuncAttributeMaxDynamicSharedMemorySize, 80 * 1024);
    (void)hipFuncSetAttribute((const void*)attn_fwd_kernel<64>, hipFuncAttributeMaxDynamicSharedMemorySize, 80 * 1024);
    attr = true;
  }
  if (a.dh == 32)
    hipLaunchKernelGGL(attn_fwd_kernel<32>, dim3(a.nz), dim3(256), smem, s, a.QKV, a.O, a.ldo, a.N, a.H);
  else
    hipLaunchKernelGGL(attn_fwd_kernel<64>, dim3(a.nz), dim3(256), smem, s, a.QKV, a.O, a.ldo, a.N, a.H);
  return 0;
}

int launch_attn_lap(const AttnArgs& a, hipStream_t s) {
  if (!attn_lap_supported(a.N, a.dh) || a.nz <= 0 || a.R > a.Rp || a.R < ROW_J0 || a.ldo % 4) return -1;
  return a.h3 ? launch_mode<true>(a, s) : launch_mode<false>(a, s);
}
