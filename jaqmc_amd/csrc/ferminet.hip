// FermiNet streaming kernels: input features, two-electron stream, spin-channel
// aggregation and the envelope, all in the tangent-row layout of common.h.
//
// Forward-Laplacian bookkeeping follows the reference's rules
// (laplacian/primitives/*.py) specialised to this graph:
//   * the atom-electron features of electron i depend only on x_i (Local1):
//     three non-zero Jacobian rows, written densely;
//   * the pair stream h2[a][b] depends on x_a and x_b only (Local2): it is kept
//     in 8 rows per pair -- value, Laplacian, d/dx_a (3), d/dx_b (3);
//   * means over electrons are linear, so they are expanded to dense rows when
//     aggregated (reductions.py:25-61).

#include <type_traits>

#include "kernels.h"

namespace {

// ---------------------------------------------------------------------------
// Layer-0 features (wavefunction/input/atomic.py:46-79, rescale=False):
//   ae[i, 4I + 0] = |x_i - R_I|,  ae[i, 4I + 1 + d] = (x_i - R_I)_d
//   ee[a, b, 0]  = |x_a - x_b| (0 on the diagonal), ee[a, b, 1 + d] = (x_a - x_b)_d
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void ae_features_kernel(FermiDims d, const float* __restrict__ elec,
                                                          const float* __restrict__ atoms, T* __restrict__ X, long ldx) {
  // One workgroup per unit (electron of a walker): the displacements to the M
  // nuclei are computed once into LDS; thread items are (row r, nucleus I),
  // each writing the four features (|d|, d) of one nucleus.
  extern __shared__ float geo[];  // [M][4]: dx, dy, dz, r
  const int u = blockIdx.x;
  const int w = walker_of_unit(u, d.Bc, d.n_up, d.n_dn), i = elec_of_unit(u, d.Bc, d.n_up, d.n_dn);
  const float* xi = elec + ((long)w * d.N + i) * 3;
  for (int I = threadIdx.x; I < d.M; I += blockDim.x) {
    const float dx = xi[0] - atoms[3 * I], dy = xi[1] - atoms[3 * I + 1], dz = xi[2] - atoms[3 * I + 2];
    geo[4 * I] = dx;
    geo[4 * I + 1] = dy;
    geo[4 * I + 2] = dz;
    geo[4 * I + 3] = sqrtf(dx * dx + dy * dy + dz * dz);
  }
  __syncthreads();
  T* Xu = X + (long)u * d.Rp * ldx;
  const int items = d.Rp * d.M;
  for (int idx = threadIdx.x; idx < items; idx += blockDim.x) {
    const int r = idx / d.M, I = idx - r * d.M;
    const float4 g = *(const float4*)&geo[4 * I];
    float o[4] = {0.f, 0.f, 0.f, 0.f};
    if (r == ROW_VAL) {
      o[0] = g.w;
      o[1] = g.x;
      o[2] = g.y;
      o[3] = g.z;
    } else if (r == ROW_LAP) {
      o[0] = 2.f / g.w;
    } else {
      const int c = r - ROW_J0;
      if (c < 3 * d.N && c / 3 == i) {
        const int ax = c - 3 * i;
        o[0] = (ax == 0 ? g.x : (ax == 1 ? g.y : g.z)) / g.w;
        o[1] = ax == 0 ? 1.f : 0.f;  // static indices keep o in registers
        o[2] = ax == 1 ? 1.f : 0.f;
        o[3] = ax == 2 ? 1.f : 0.f;
      }
    }
    T* dst = Xu + (long)r * ldx + 4 * I;
    if constexpr (std::is_same<T, float>::value) {
      *(float4*)dst = *(float4*)o;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) dst[e] = from_f<T>(o[e]);
    }
  }
}

// Pair-stream input, H2[w][a][b][row][4], RP2 = 8 rows (Laplacian) or 1.
__global__ void ee_features_kernel(FermiDims d, int RP2, const float* __restrict__ elec, float* __restrict__ H2) {
  const long total = (long)d.Bc * d.N * d.N;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < total; p += (long)gridDim.x * blockDim.x) {
    const int b = p % d.N;
    const int a = (p / d.N) % d.N;
    const int w = p / ((long)d.N * d.N);
    float* o = H2 + p * RP2 * 4;
    const float* xa = elec + ((long)w * d.N + a) * 3;
    const float* xb = elec + ((long)w * d.N + b) * 3;
    const float dv[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
    if (a == b) {
      for (int q = 0; q < RP2 * 4; ++q) o[q] = 0.f;
      continue;
    }
    const float rr = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    o[0] = rr;
    o[1] = dv[0];
    o[2] = dv[1];
    o[3] = dv[2];
    if (RP2 == 1) continue;
    // Laplacian over both particles: 2 * (2/r) for r, 0 for displacements.
    o[4] = 4.f / rr;
    o[5] = o[6] = o[7] = 0.f;
    for (int ax = 0; ax < 3; ++ax) {
      float* ja = o + (2 + ax) * 4;
      float* jb = o + (5 + ax) * 4;
      ja[0] = dv[ax] / rr;
      jb[0] = -dv[ax] / rr;
      for (int e = 0; e < 3; ++e) {
        ja[1 + e] = e == ax ? 1.f : 0.f;
        jb[1 + e] = e == ax ? -1.f : 0.f;
      }
    }
  }
}

// One two-electron layer: out = tanh(in W + b) (+ (in + out)/sqrt2 when
// widths match), FermiLayers._update_double_stream (ferminet.py:57-63).
// Thread = one tangent row of one pair (RP2 = 8 rows: value, Laplacian,
// d/dx_a, d/dx_b); W staged in LDS and read as broadcast float4; the Local2
// tanh rule sums J^2 over the pair's six Jacobian rows with DPP lane moves
// inside the 8-lane group of the pair.
template <int RP2, int FIN, int FOUT>
__global__ __launch_bounds__(256) void pair_layer_kernel(long npairs, const float* __restrict__ in,
                                                         float* __restrict__ out, const float* __restrict__ W,
                                                         const float* __restrict__ bias, int residual) {
  __shared__ __attribute__((aligned(16))) float Ws[FIN * FOUT];
  __shared__ float bs[FOUT];
  for (int idx = threadIdx.x; idx < FIN * FOUT; idx += 256) Ws[idx] = W[idx];
  for (int idx = threadIdx.x; idx < FOUT; idx += 256) bs[idx] = bias[idx];
  __syncthreads();
  const long rowid = (long)blockIdx.x * 256 + threadIdx.x;  // pair * RP2 + r
  const long nrows = npairs * RP2;
  const bool live = rowid < nrows;
  const int r = (int)(rowid % RP2);
  const float* xp = in + (live ? rowid : 0) * FIN;
  float acc[FOUT];
#pragma unroll
  for (int f = 0; f < FOUT; ++f) acc[f] = (r == 0) ? bs[f] : 0.f;
  if constexpr (FIN <= 4) {
    const float4 xv = *(const float4*)xp;
    const float xk[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll 2
    for (int k = 0; k < FIN; ++k)
#pragma unroll
      for (int f = 0; f < FOUT; f += 4) {
        const float4 w = *(const float4*)&Ws[k * FOUT + f];
        acc[f] += xk[k] * w.x;
        acc[f + 1] += xk[k] * w.y;
        acc[f + 2] += xk[k] * w.z;
        acc[f + 3] += xk[k] * w.w;
      }
  } else {
    // the input row is consumed four features at a time (16-byte loads; the
    // row stays in L1), so no private array is indexed with a runtime k
#pragma unroll 2
    for (int k4 = 0; k4 < FIN; k4 += 4) {
      const float4 xv = *(const float4*)(xp + k4);
      const float xk[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int f = 0; f < FOUT; f += 4) {
          const float4 w = *(const float4*)&Ws[(k4 + kk) * FOUT + f];
          acc[f] += xk[kk] * w.x;
          acc[f + 1] += xk[kk] * w.y;
          acc[f + 2] += xk[kk] * w.z;
          acc[f + 3] += xk[kk] * w.w;
        }
    }
  }
  float* yp = out + (live ? rowid : 0) * FOUT;
  static_assert(RP2 == 1 || RP2 == 8, "pair rows");
  const bool upper = (threadIdx.x & 4) != 0;
#pragma unroll
  for (int f = 0; f < FOUT; ++f) {
    // the pair's 8 rows are lanes 8g .. 8g+7: sum J^2 over rows 2..7 and
    // broadcast the value row with DPP inside the group (quad_perm xor 1,
    // xor 2, then the half-row mirror joins the two quads)
    float sq = (r >= 2) ? acc[f] * acc[f] : 0.f;
    float v = acc[f];
    if (RP2 > 1) {
      sq += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, sq), 0xB1, 0xF, 0xF, false));
      sq += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, sq), 0x4E, 0xF, 0xF, false));
      sq += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, sq), 0x141, 0xF, 0xF, false));
      const float b = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x00, 0xF, 0xF, false));
      const float m = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, b), 0x141, 0xF, 0xF, false));
      v = upper ? m : b;
    }
    const float t = tanhf(v);
    const float d1 = 1.f - t * t;
    float o;
    if (r == 0)
      o = t;
    else if (r == 1)
      o = d1 * acc[f] - 2.f * t * d1 * sq;
    else
      o = d1 * acc[f];
    acc[f] = o;
  }
  if (residual && FIN == FOUT) {
#pragma unroll
    for (int f = 0; f < FOUT; f += 4) {
      const float4 xv = *(const float4*)(xp + f);
      acc[f] = (acc[f] + xv.x) * 0.70710678118654752f;
      acc[f + 1] = (acc[f + 1] + xv.y) * 0.70710678118654752f;
      acc[f + 2] = (acc[f + 2] + xv.z) * 0.70710678118654752f;
      acc[f + 3] = (acc[f + 3] + xv.w) * 0.70710678118654752f;
    }
  }
  if (live) {
#pragma unroll
    for (int f = 0; f < FOUT; f += 4) *(float4*)(yp + f) = *(float4*)&acc[f];
  }
}

// The FIN -> 32 two-electron layer (FIN = 4 or 32) on v_mfma_f32_16x16x4f32:
// one wavefront per block of 16 rows (two pairs of eight tangent rows, or 16
// value rows).  With KK = FIN / 4 steps, lane (r = lane%16, g = lane/16)
// supplies A = in[row r][KK g + kk] and B = W[KK g + kk][16 nb + r] at step kk
// (the k order inside an MFMA step is a
// permutation both operands share); result element v is row 4g + v, column
// 16 nb + r, so a pair's rows 0..3 / 4..7 sit in lane groups 2p / 2p + 1 and
// the tanh rule's sums over the Jacobian rows are an in-thread sum plus one
// exchange with lane r ^ 16.  Same rule and residual as pair_layer_kernel.
template <int RP2, int FIN>
__global__ __launch_bounds__(256) void pair_layer32_mfma_kernel(long npairs, const float* __restrict__ in,
                                                                float* __restrict__ out, const float* __restrict__ W,
                                                                const float* __restrict__ bias, int residual) {
  // per-wave 16 x 32 output tile (row stride 33), written in the MFMA layout
  // and read back as 32-byte row segments for the stores
  __shared__ float tile[4][16 * 33];
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4, wave = threadIdx.x >> 6;
  const long nrows = npairs * RP2;
  const long blk = (long)blockIdx.x * 4 + wave;
  const long row0 = blk * 16;  // blocks past the end compute on clamped rows and store nothing
  static_assert(FIN == 4 || FIN == 32, "input widths 4 and 32");
  constexpr int KK = FIN / 4;
  float wb[2][KK];
#pragma unroll
  for (int nb = 0; nb < 2; ++nb)
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) wb[nb][kk] = W[(KK * g + kk) * 32 + 16 * nb + r];
  const long ra = row0 + r < nrows ? row0 + r : nrows - 1;
  float xa[KK];
  if constexpr (FIN == 32) {
    const float4 x0 = *(const float4*)(in + ra * 32 + 8 * g);
    const float4 x1 = *(const float4*)(in + ra * 32 + 8 * g + 4);
    xa[0] = x0.x; xa[1] = x0.y; xa[2] = x0.z; xa[3] = x0.w;
    xa[4] = x1.x; xa[5] = x1.y; xa[6] = x1.z; xa[7] = x1.w;
  } else {
    xa[0] = in[ra * 4 + g];
  }
  f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
  for (int kk = 0; kk < KK; ++kk)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[kk], wb[nb][kk], acc[nb], 0, 0, 0);
  const float r2 = 0.70710678118654752f;
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    const int col = 16 * nb + r;
    const float bv = bias[col];
    float o[4];
    if constexpr (RP2 == 8) {
      const bool lo = (g & 1) == 0;  // rows 0..3 of the pair (value, Laplacian, first Jacobian rows)
      if (lo) acc[nb][0] += bv;
      float sq = 0.f;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (!lo || v >= 2) sq += acc[nb][v] * acc[nb][v];
      sq += __shfl_xor(sq, 16);
      const float y0 = __shfl(acc[nb][0], lane & ~16);
      const float t = tanhf(y0);
      const float d1 = 1.f - t * t;
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = d1 * acc[nb][v];
      if (lo) {
        o[0] = t;
        o[1] = d1 * acc[nb][1] - 2.f * t * d1 * sq;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = tanhf(acc[nb][v] + bv);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) tile[wave][(4 * g + v) * 33 + col] = o[v];
  }
  __syncthreads();
  // lane: row lane/4 of the block, columns 8 (lane%4) .. + 7
  const int tr = lane >> 2, tc = 8 * (lane & 3);
  const long row = row0 + tr;
  if (row < nrows) {
    float y[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) y[e] = tile[wave][tr * 33 + tc + e];
    if (FIN == 32 && residual) {
      const float4 a = *(const float4*)(in + row * 32 + tc), b = *(const float4*)(in + row * 32 + tc + 4);
      const float xr[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int e = 0; e < 8; ++e) y[e] = (y[e] + xr[e]) * r2;
    }
    *(float4*)(out + row * 32 + tc) = float4{y[0], y[1], y[2], y[3]};
    *(float4*)(out + row * 32 + tc + 4) = float4{y[4], y[5], y[6], y[7]};
  }
}

// Generic fallback (any widths): thread = (pair, output feature).
template <int RP2>
__global__ void pair_layer_generic_kernel(long npairs, const float* __restrict__ in, int fin, float* __restrict__ out,
                                          int fout, const float* __restrict__ W, const float* __restrict__ bias,
                                          int residual) {
  const long total = npairs * fout;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int f = idx % fout;
    const long p = idx / fout;
    const float* x = in + p * RP2 * fin;
    float acc[RP2];
#pragma unroll
    for (int r = 0; r < RP2; ++r) acc[r] = 0.f;
    for (int k = 0; k < fin; ++k) {
      const float wk = W[(long)k * fout + f];
#pragma unroll
      for (int r = 0; r < RP2; ++r) acc[r] += x[r * fin + k] * wk;
    }
    acc[0] += bias[f];
    float* y = out + p * RP2 * fout;
    const float t = tanhf(acc[0]);
    const float d1 = 1.f - t * t;
    float o[RP2];
    o[0] = t;
    if constexpr (RP2 > 1) {
      float sq = 0.f;
#pragma unroll
      for (int r = 2; r < RP2; ++r) sq += acc[r] * acc[r];
      o[1] = d1 * acc[1] - 2.f * t * d1 * sq;
#pragma unroll
      for (int r = 2; r < RP2; ++r) o[r] = d1 * acc[r];
    }
#pragma unroll
    for (int r = 0; r < RP2; ++r) {
      float v = o[r];
      if (residual) v = (v + x[r * fin + f]) * 0.70710678118654752f;
      y[r * fout + f] = v;
    }
  }
}

// g_two of FermiLayers.aggregate_features (ferminet.py:86-89): for electron i and
// channel s, m_s[i] = mean_{j in s} h2[j][i].  Expanded to dense rows:
//   d/dx_i   <- mean_j (d/dx_b rows of pair (j, i))
//   d/dx_j   <- (d/dx_a rows of pair (j, i)) / n_s,   j in s
// One workgroup per (walker, electron) stages the column H2[w][:, i] in LDS
// and writes the unit's whole block X[unit][0:Rp][col0:ldx) row by row;
// columns past nch*f2 are zero-filled (GEMM K padding).  NTH threads: 1024
// when the staged column is large enough to leave room for few workgroups
// per CU (large N), else 256.
template <typename T, int NTH>
__global__ __launch_bounds__(NTH) void pair_means_kernel(FermiDims d, int RP2, const float* __restrict__ H2, int f2,
                                                         T* __restrict__ X, long ldx, int col0) {
  extern __shared__ float col[];  // [j][RP2][f2]
  float* part = col + d.N * RP2 * f2;  // [s][5][f2]: sums of value, lap, d/dx_i (3)
  const int w = blockIdx.x / d.N, i = blockIdx.x % d.N;
  const int tid = threadIdx.x;
  const long pstride = (long)RP2 * f2;
  const float* base = H2 + ((long)w * d.N * d.N + i) * pstride;  // pair (0, i)
  if ((RP2 * f2) % 4 == 0) {
    // 16-byte loads, eight in flight per thread
    const int q4 = RP2 * f2 / 4;
#pragma unroll 8
    for (int idx = tid; idx < d.N * q4; idx += NTH) {
      const int j = idx / q4, q = idx - j * q4;
      *(float4*)&col[4 * idx] = *(const float4*)(base + (long)j * d.N * pstride + 4 * q);
    }
  } else {
#pragma unroll 8
    for (int idx = tid; idx < d.N * RP2 * f2; idx += NTH) {
      const int j = idx / (RP2 * f2), q = idx % (RP2 * f2);
      col[idx] = base[(long)j * d.N * pstride + q];
    }
  }
  __syncthreads();
  const int nrow_sum = RP2 > 1 ? 5 : 1;
  for (int idx = tid; idx < d.nch * nrow_sum * f2; idx += NTH) {
    const int s = idx / (nrow_sum * f2), q = (idx / f2) % nrow_sum, f = idx % f2;
    const int j0 = d.nch == 1 ? 0 : (s == 0 ? 0 : d.n_up);
    const int ns = d.nch == 1 ? d.N : (s == 0 ? d.n_up : d.n_dn);
    const int row = q < 2 ? q : 5 + (q - 2);  // value, lap, d/dx_b (rows 5..7)
    float acc = 0.f;
    for (int j = j0; j < j0 + ns; ++j) acc += col[(j * RP2 + row) * f2 + f];
    part[(s * 5 + q) * f2 + f] = acc / ns;
  }
  __syncthreads();
  const int u = unit_of(w, i, d.Bc, d.n_up, d.n_dn);
  const int ncol = (int)(ldx - col0);
  T* xu = X + (long)u * d.Rp * ldx + col0;
  // lanes walk columns (per-column constants hoisted), the waves walk rows
  for (int cc = tid & 63; cc < ncol; cc += 64) {
    const bool live = cc < d.nch * f2;
    const int s = live ? cc / f2 : 0, f = live ? cc - s * f2 : 0;
    const int j0 = d.nch == 1 ? 0 : (s == 0 ? 0 : d.n_up);
    const int ns = d.nch == 1 ? d.N : (s == 0 ? d.n_up : d.n_dn);
    const float inv = 1.f / ns;
    for (int r = tid >> 6; r < d.Rp; r += NTH / 64) {
      float out = 0.f;
      if (live) {
        if (r == ROW_VAL || r == ROW_LAP) {
          out = part[(s * 5 + r) * f2 + f];
        } else if (r - ROW_J0 < 3 * d.N) {
          const int c = r - ROW_J0;
          const int e = c / 3, ax = c - 3 * e;
          if (e == i)
            out = part[(s * 5 + 2 + ax) * f2 + f];
          else if (e >= j0 && e < j0 + ns)
            out = col[(e * RP2 + 2 + ax) * f2 + f] * inv;
        }
      }
      xu[(long)r * ldx + cc] = from_f<T>(out);
    }
  }
}

// Spin-channel means of the one-electron stream, rsplit workgroups per
// (walker, channel), each over a slice of the rows: g[w][r][s*f1 + f] = mean_{i in s} X[unit(w, i)][r][f]
// with float4 loads and stores (f1 % 4 == 0); columns >= nch*f1 of g are
// zero-filled by the s = nch-1 workgroups, each for its rows.
__global__ __launch_bounds__(256) void spin_means4_kernel(FermiDims d, int rsplit, const float* __restrict__ X,
                                                          long ldx, int f1, float* __restrict__ g, long ldg) {
  const int ws = blockIdx.x / rsplit, rs = blockIdx.x % rsplit;
  const int w = ws / d.nch, s = ws % d.nch;
  const int rows = (d.Rp + rsplit - 1) / rsplit, r0 = rs * rows;
  const int nr = d.Rp - r0 < rows ? d.Rp - r0 : rows;  // <= 0: nothing to do
  const int i0 = d.nch == 1 ? 0 : (s == 0 ? 0 : d.n_up);
  const int ns = d.nch == 1 ? d.N : (s == 0 ? d.n_up : d.n_dn);
  const int u0 = unit_of(w, i0, d.Bc, d.n_up, d.n_dn);  // units of (w, s) are consecutive
  const float inv = 1.f / ns;
  const int q4 = f1 / 4;
  for (int idx = threadIdx.x; idx < nr * q4; idx += 256) {
    const int r = r0 + idx / q4, f = (idx % q4) * 4;
    float4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int k = 0; k < ns; ++k) {
      const float4 v = *(const float4*)(X + ((long)(u0 + k) * d.Rp + r) * ldx + f);
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    acc.x *= inv;
    acc.y *= inv;
    acc.z *= inv;
    acc.w *= inv;
    *(float4*)(g + ((long)w * d.Rp + r) * ldg + s * f1 + f) = acc;
  }
  if (s == d.nch - 1) {
    const int pad = (int)(ldg - d.nch * f1);
    for (int idx = threadIdx.x; idx < nr * pad; idx += 256) {
      const int r = r0 + idx / pad, c = idx % pad;
      g[((long)w * d.Rp + r) * ldg + d.nch * f1 + c] = 0.f;
    }
  }
}

// g_one: spin-channel means of the one-electron stream, per walker and row:
// g[w][r][s*f1 + f] = mean_{i in s} X[unit(w, i)][r][f]; columns >= nch*f1 zeroed.
template <typename T>
__global__ void spin_means_kernel(FermiDims d, const T* __restrict__ X, long ldx, int f1, T* __restrict__ g,
                                  long ldg) {
  const long total = (long)d.Bc * d.Rp * ldg;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int cc = idx % ldg;
    const long wr = idx / ldg;
    const int r = wr % d.Rp;
    const int w = wr / d.Rp;
    float out = 0.f;
    if (cc < d.nch * f1) {
      const int s = cc / f1, f = cc % f1;
      const int i0 = d.nch == 1 ? 0 : (s == 0 ? 0 : d.n_up);
      const int ns = d.nch == 1 ? d.N : (s == 0 ? d.n_up : d.n_dn);
      float acc = 0.f;
      for (int i = i0; i < i0 + ns; ++i) {
        const int u = unit_of(w, i, d.Bc, d.n_up, d.n_dn);
        acc += to_f<T>(X[(long)u * d.Rp * ldx + (long)r * ldx + f]);
      }
      out = acc / ns;
    }
    g[idx] = from_f<T>(out);
  }
}

// Envelope (wavefunction/output/envelope.py:111-140), per electron i of spin
// channel s, det k, orbital o:  e = sum_I pi[o,I,k] exp(-a r_iI), a = |sigma| or
// sigma.  Gradient wrt x_i: sum_I -a f (x_i - R_I)/r; Laplacian: sum_I f (a^2 - 2a/r).
// Output env[w][i][k*N + o][5] = (value, Laplacian, d/dx, d/dy, d/dz) in
// Laplacian mode, env[w][i][k*N + o] = value in forward mode.
// One workgroup per electron index i and block of EW walkers (one spin
// channel, so one parameter set): the EW electrons' displacements to the
// nuclei are computed once into LDS, and the threads walk (k, o) with the
// parameters laid out [I][k][o] (coalesced across the threads for each
// nucleus), each parameter pair loaded once for the EW electrons.
constexpr int kEnvWalkers = 8;

template <bool LAP>
__global__ __launch_bounds__(256) void envelope_kernel(FermiDims d, const float* __restrict__ elec,
                                                       const float* __restrict__ atoms, const float* __restrict__ pi0,
                                                       const float* __restrict__ sg0, const float* __restrict__ pi1,
                                                       const float* __restrict__ sg1, int kind,
                                                       float* __restrict__ env) {
  constexpr int EW = kEnvWalkers;
  extern __shared__ float geo[];  // [EW][M][4]: dx, dy, dz, r log2(e); then [EW][M]: 1 / r
  const int i = blockIdx.x, w0 = blockIdx.y * EW;
  const int ne = min(EW, d.Bc - w0);
  const int KN = d.K * d.N, M = d.M;
  constexpr int ES = LAP ? 5 : 1;  // env_stride
  if (kind == 2) {  // null envelope
    for (int e = 0; e < ne; ++e) {
      float* out = env + ((long)(w0 + e) * d.N + i) * KN * ES;
      for (int ko = threadIdx.x; ko < KN; ko += blockDim.x) {
        out[ko * ES] = 1.f;
        if constexpr (LAP) out[ko * 5 + 1] = out[ko * 5 + 2] = out[ko * 5 + 3] = out[ko * 5 + 4] = 0.f;
      }
    }
    return;
  }
  float* const irs = geo + 4 * EW * M;
  for (int t = threadIdx.x; t < EW * M; t += blockDim.x) {
    const int e = t / M, I = t % M;
    const float* xi = elec + ((long)(w0 + (e < ne ? e : 0)) * d.N + i) * 3;  // past the last walker: a copy
    const float dx = xi[0] - atoms[3 * I], dy = xi[1] - atoms[3 * I + 1], dz = xi[2] - atoms[3 * I + 2];
    geo[4 * t] = dx;
    geo[4 * t + 1] = dy;
    geo[4 * t + 2] = dz;
    const float r = sqrtf(dx * dx + dy * dy + dz * dz);
    geo[4 * t + 3] = r * 1.4426950408889634f;  // r log2(e): exp(-a r) = 2^(-a r log2 e)
    irs[t] = 1.f / r;
  }
  __syncthreads();
  const bool up = d.nch == 1 || i < d.n_up;
  const float* pi = up ? pi0 : pi1;
  const float* sg = up ? sg0 : sg1;
  for (int ko = threadIdx.x; ko < KN; ko += blockDim.x) {
    float v[EW], l[EW], gx[EW], gy[EW], gz[EW];
#pragma unroll
    for (int e = 0; e < EW; ++e) v[e] = l[e] = gx[e] = gy[e] = gz[e] = 0.f;
    for (int I = 0; I < M; ++I) {
      const float a = kind == 1 ? fabsf(sg[(long)I * KN + ko]) : sg[(long)I * KN + ko];
      const float p = pi[(long)I * KN + ko];
#pragma unroll
      for (int e = 0; e < EW; ++e) {
        const float4 g = *(const float4*)&geo[4 * (e * M + I)];
        const float f = p * __builtin_amdgcn_exp2f(-a * g.w);
        v[e] += f;
        if constexpr (LAP) {
          const float ir = irs[e * M + I];  // one division per atom and electron
          l[e] += f * a * (a - 2.f * ir);
          const float gs = -a * f * ir;
          gx[e] += gs * g.x;
          gy[e] += gs * g.y;
          gz[e] += gs * g.z;
        }
      }
    }
#pragma unroll
    for (int e = 0; e < EW; ++e) {
      if (e >= ne) break;
      float* out = env + ((long)(w0 + e) * d.N + i) * KN * ES;
      out[ko * ES] = v[e];
      if constexpr (LAP) {
        out[ko * 5 + 1] = l[e];
        out[ko * 5 + 2] = gx[e];
        out[ko * 5 + 3] = gy[e];
        out[ko * 5 + 4] = gz[e];
      }
    }
  }
}

inline int grid_for(long total) {
  long g = (total + 255) / 256;
  if (g > 65536L * 8) g = 65536L * 8;
  return (int)(g < 1 ? 1 : g);
}

// Forward mode (one row per pair and per unit): thread per output column of
// a unit, X[unit(w, i)][col0 + s*f2 + f] = mean_{j in s} H2[w][j][i][f]
// (summed in the same j order as the Laplacian-mode kernel), columns past
// nch*f2 zero.  No staging: a unit's work is a few hundred floats.
template <typename T>
__global__ __launch_bounds__(256) void pair_means_fwd_kernel(FermiDims d, const float* __restrict__ H2, int f2,
                                                             T* __restrict__ X, long ldx, int col0) {
  const int ncol = (int)(ldx - col0);
  const long total = (long)d.Bc * d.N * ncol;
  for (long idx = blockIdx.x * 256L + threadIdx.x; idx < total; idx += (long)gridDim.x * 256) {
    const int cc = (int)(idx % ncol);
    const long wi = idx / ncol;
    const int w = (int)(wi / d.N), i = (int)(wi % d.N);
    float out = 0.f;
    if (cc < d.nch * f2) {
      const int s = cc / f2, f = cc - s * f2;
      const int j0 = d.nch == 1 ? 0 : (s == 0 ? 0 : d.n_up);
      const int ns = d.nch == 1 ? d.N : (s == 0 ? d.n_up : d.n_dn);
      const float* p = H2 + (((long)w * d.N + j0) * d.N + i) * f2 + f;
      float acc = 0.f;
      for (int j = 0; j < ns; ++j) acc += p[(long)j * d.N * f2];
      out = acc / ns;
    }
    X[(long)unit_of(w, i, d.Bc, d.n_up, d.n_dn) * ldx + col0 + cc] = from_f<T>(out);
  }
}

}  // namespace

int launch_ferminet_features(const FermiDims& d, const float* elec, const float* atoms, void* X, long ldx,
                             bool x_bf16, float* H2, hipStream_t s) {
  const unsigned units = (unsigned)(d.Bc * d.N);
  const size_t smem = (size_t)4 * d.M * sizeof(float);
  if (units > 0 && smem <= 64 * 1024) {
    if (x_bf16)
      hipLaunchKernelGGL(ae_features_kernel<bf16>, dim3(units), 256, smem, s, d, elec, atoms, (bf16*)X, ldx);
    else
      hipLaunchKernelGGL(ae_features_kernel<float>, dim3(units), 256, smem, s, d, elec, atoms, (float*)X, ldx);
  }
  const long t2 = (long)d.Bc * d.N * d.N;
  hipLaunchKernelGGL(ee_features_kernel, grid_for(t2), 256, 0, s, d, d.Rp > 1 ? 8 : 1, elec, H2);
  return 0;
}

template <int RP2>
static bool pair_layer_fast(long npairs, const float* in, int fin, float* out, int fout, const float* W,
                            const float* b, int res, hipStream_t s) {
  const unsigned grid = (unsigned)((npairs * RP2 + 255) / 256);
#define JQ_PL(FI, FO)                                                                                     \
  if (fin == FI && fout == FO) {                                                                          \
    hipLaunchKernelGGL((pair_layer_kernel<RP2, FI, FO>), dim3(grid), dim3(256), 0, s, npairs, in, out, W, b, \
                       res);                                                                              \
    return true;                                                                                          \
  }
  if ((fin == 32 || fin == 4) && fout == 32) {
    const long nblk = (npairs * RP2 + 15) / 16;
    const dim3 grid32((unsigned)((nblk + 3) / 4));
    if (fin == 32)
      hipLaunchKernelGGL((pair_layer32_mfma_kernel<RP2, 32>), grid32, dim3(256), 0, s, npairs, in, out, W, b, res);
    else
      hipLaunchKernelGGL((pair_layer32_mfma_kernel<RP2, 4>), grid32, dim3(256), 0, s, npairs, in, out, W, b, res);
    return true;
  }
  JQ_PL(4, 16)
  JQ_PL(16, 16)
#undef JQ_PL
  return false;
}

int launch_pair_layer(const FermiDims& d, const float* H2in, int fin, float* H2out, int fout, const float* W,
                      const float* b, bool residual, hipStream_t s) {
  const long npairs = (long)d.Bc * d.N * d.N;
  const long total = npairs * fout;
  const int res = residual ? 1 : 0;
  if (d.Rp > 1) {
    if (!pair_layer_fast<8>(npairs, H2in, fin, H2out, fout, W, b, res, s))
      hipLaunchKernelGGL(pair_layer_generic_kernel<8>, grid_for(total), 256, 0, s, npairs, H2in, fin, H2out, fout, W,
                         b, res);
  } else {
    if (!pair_layer_fast<1>(npairs, H2in, fin, H2out, fout, W, b, res, s))
      hipLaunchKernelGGL(pair_layer_generic_kernel<1>, grid_for(total), 256, 0, s, npairs, H2in, fin, H2out, fout, W,
                         b, res);
  }
  return 0;
}

int launch_pair_means(const FermiDims& d, const float* H2, int f2, void* X, long ldx, int col0, bool x_bf16,
                      hipStream_t s) {
  const int RP2 = d.Rp > 1 ? 8 : 1;
  if (RP2 == 1) {
    const long total = (long)d.Bc * d.N * (ldx - col0);
    if (total <= 0) return 0;
    if (x_bf16)
      hipLaunchKernelGGL(pair_means_fwd_kernel<bf16>, grid_for(total), 256, 0, s, d, H2, f2, (bf16*)X, ldx, col0);
    else
      hipLaunchKernelGGL(pair_means_fwd_kernel<float>, grid_for(total), 256, 0, s, d, H2, f2, (float*)X, ldx, col0);
    return 0;
  }
  const size_t smem = ((size_t)d.N * RP2 * f2 + 2 * 5 * f2) * 4;
  if (smem > 64 * 1024) return -1;
  const dim3 grid((unsigned)((long)d.Bc * d.N));
  if (smem > 40 * 1024) {
    if (x_bf16)
      hipLaunchKernelGGL((pair_means_kernel<bf16, 1024>), grid, 1024, smem, s, d, RP2, H2, f2, (bf16*)X, ldx, col0);
    else
      hipLaunchKernelGGL((pair_means_kernel<float, 1024>), grid, 1024, smem, s, d, RP2, H2, f2, (float*)X, ldx, col0);
  } else {
    if (x_bf16)
      hipLaunchKernelGGL((pair_means_kernel<bf16, 256>), grid, 256, smem, s, d, RP2, H2, f2, (bf16*)X, ldx, col0);
    else
      hipLaunchKernelGGL((pair_means_kernel<float, 256>), grid, 256, smem, s, d, RP2, H2, f2, (float*)X, ldx, col0);
  }
  return 0;
}

int launch_spin_means(const FermiDims& d, const void* X, long ldx, int f1, void* g, long ldg, bool bf16io,
                      hipStream_t s) {
  if (!bf16io && f1 % 4 == 0 && ldx % 4 == 0 && ldg % 4 == 0) {
    // rows of a (walker, channel) split over several workgroups when there
    // are few walkers, so the launch has a few thousand workgroups
    const long wg = (long)d.Bc * d.nch;
    int rsplit = (int)((4096 + wg - 1) / wg);
    const int rmax = d.Rp >= 16 ? d.Rp / 8 : 1;
    if (rsplit > rmax) rsplit = rmax;
    hipLaunchKernelGGL(spin_means4_kernel, dim3((unsigned)(wg * rsplit)), 256, 0, s, d, rsplit, (const float*)X, ldx,
                       f1, (float*)g, ldg);
    return 0;
  }
  const long total = (long)d.Bc * d.Rp * ldg;
  if (bf16io)
    hipLaunchKernelGGL(spin_means_kernel<bf16>, grid_for(total), 256, 0, s, d, (const bf16*)X, ldx, f1, (bf16*)g,
                       ldg);
  else
    hipLaunchKernelGGL(spin_means_kernel<float>, grid_for(total), 256, 0, s, d, (const float*)X, ldx, f1,
                       (float*)g, ldg);
  return 0;
}

int launch_envelope(const FermiDims& d, const float* elec, const float* atoms, const float* const* pi,
                    const float* const* sigma, int env_kind, float* env, hipStream_t s) {
  if ((long)d.Bc * d.N <= 0) return 0;
  const size_t smem = (size_t)5 * kEnvWalkers * d.M * sizeof(float);
  if (smem > 64 * 1024) return -1;
  const dim3 grid((unsigned)d.N, (unsigned)((d.Bc + kEnvWalkers - 1) / kEnvWalkers));
  if (env_stride(d) == 5)
    hipLaunchKernelGGL(envelope_kernel<true>, grid, 256, smem, s, d, elec, atoms, pi[0], sigma[0],
                       d.nch > 1 ? pi[1] : pi[0], d.nch > 1 ? sigma[1] : sigma[0], env_kind, env);
  else
    hipLaunchKernelGGL(envelope_kernel<false>, grid, 256, smem, s, d, elec, atoms, pi[0], sigma[0],
                       d.nch > 1 ? pi[1] : pi[0], d.nch > 1 ? sigma[1] : sigma[0], env_kind, env);
  return 0;
}
