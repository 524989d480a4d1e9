// Psiformer streaming kernels in the tangent-row layout (walker-major units):
// rescaled input features, LayerNorm and the attention softmax, each with its
// forward-Laplacian rule.
//
// Attention scores and outputs are batched MFMA GEMMs (gemm.hip); this file
// holds the parts with reductions over features or keys.

#include "kernels.h"

namespace {

inline int grid_for(long total) {
  long g = (total + 255) / 256;
  if (g > 65536L * 8) g = 65536L * 8;
  return (int)(g < 1 ? 1 : g);
}

// MoleculeFeatures(rescale=True) (wavefunction/input/atomic.py:56-63) plus the
// spin column of PsiformerBackbone (backbone/psiformer.py:165-172):
//   f[4I]     = g = log(1 + r_iI)
//   f[4I+1+a] = d_a g / r                (d = x_i - R_I)
//   f[4M]     = +1 (spin up) / -1 (spin down)
// Derivatives wrt x_i only (Local1): with q = g / r,
//   grad g = g' d / r,  lap g = g'' + 2 g'/r
//   d/dx_b (d_a q) = delta_ab q + d_a q' d_b / r,  lap (d_a q) = d_a (q'' + 2 q'/r) + 2 q' d_a / r.
// With rescale = 0 (MoleculeFeatures(rescale=False)) the features are
// [r, d] as in FermiNet: grad r = d / r, lap r = 2 / r; grad d_a = e_a, lap d_a = 0.
__global__ void psi_features_kernel(FermiDims d, const float* __restrict__ elec, const float* __restrict__ atoms,
                                    float* __restrict__ X, int kin, int rescale) {
  const long total = (long)d.Bc * d.N * d.Rp * kin;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int f = idx % kin;
    const long ur = idx / kin;
    const int r = ur % d.Rp;
    const int u = ur / d.Rp;
    const int w = u / d.N, i = u % d.N;
    float out = 0.f;
    // rows other than the value, the Laplacian and electron i's own three
    // Jacobian rows are zero: whole waves of them skip the geometry
    const bool live = r == ROW_VAL || r == ROW_LAP || (r - ROW_J0 < 3 * d.N && (r - ROW_J0) / 3 == i);
    if (f < 4 * d.M) {
      if (!live) {
        X[ur * kin + f] = 0.f;
        continue;
      }
      const int I = f >> 2, e = f & 3;
      const float* xi = elec + ((long)w * d.N + i) * 3;
      const float dv[3] = {xi[0] - atoms[3 * I], xi[1] - atoms[3 * I + 1], xi[2] - atoms[3 * I + 2]};
      const float rr = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
      const float g = log1pf(rr), g1 = 1.f / (1.f + rr), g2 = -g1 * g1;
      if (!rescale) {
        const int b = (r - ROW_J0) % 3;
        if (e == 0) {
          if (r == ROW_VAL)
            out = rr;
          else if (r == ROW_LAP)
            out = 2.f / rr;
          else
            out = dv[b] / rr;
        } else {
          if (r == ROW_VAL)
            out = dv[e - 1];
          else if (r != ROW_LAP)
            out = b == e - 1 ? 1.f : 0.f;
        }
      } else if (e == 0) {
        if (r == ROW_VAL)
          out = g;
        else if (r == ROW_LAP)
          out = g2 + 2.f * g1 / rr;
        else if (r - ROW_J0 < 3 * d.N && (r - ROW_J0) / 3 == i)
          out = g1 * dv[(r - ROW_J0) % 3] / rr;
      } else {
        const int ax = e - 1;
        const float q = g / rr;
        const float q1 = g1 / rr - g / (rr * rr);
        const float q2 = g2 / rr - 2.f * g1 / (rr * rr) + 2.f * g / (rr * rr * rr);
        if (r == ROW_VAL)
          out = dv[ax] * q;
        else if (r == ROW_LAP)
          out = dv[ax] * (q2 + 2.f * q1 / rr) + 2.f * q1 * dv[ax] / rr;
        else if (r - ROW_J0 < 3 * d.N && (r - ROW_J0) / 3 == i) {
          const int b = (r - ROW_J0) % 3;
          out = (b == ax ? q : 0.f) + dv[ax] * q1 * dv[b] / rr;
        }
      }
    } else if (f == 4 * d.M && r == ROW_VAL) {
      out = i < d.n_up ? 1.f : -1.f;
    }
    X[ur * kin + f] = out;
  }
}

// LayerNorm forward Laplacian (flax nn.LayerNorm, epsilon 1e-5;
// backbone/psiformer.py:74,85).  One workgroup per unit streams the unit's
// Rp x D tangent tile from L2: thread r owns tangent row r for the row
// statistics (mean, <d_r, d_0>, |d_r|^2), then thread f owns feature f for
// the output and the cross term.  With d = x - mean(x) (linear),
// v = mean(d0^2), s = (v + eps)^-1/2:
//   J_v = 2/D sum_f d0 J_d,     lap_v = 2/D sum_f (d0 lap_d + sum_c J_d^2)
//   J_s = s' J_v,                lap_s = s' lap_v + s'' sum_c J_v^2
//   y = scale (d s) + bias:  J_y = scale (J_d s + d0 J_s),
//   lap_y = scale (lap_d s + d0 lap_s + 2 sum_c J_d J_s)
template <int D>
__global__ __launch_bounds__(256) void layernorm_lap_kernel(int Rp, int R, const float* __restrict__ x,
                                                            float* __restrict__ y, const float* __restrict__ scale,
                                                            const float* __restrict__ bias, float eps) {
  __shared__ float d0s[D];
  __shared__ float mu[256], js[256], red[8];
  const long base = (long)blockIdx.x * Rp * D;
  const int tid = threadIdx.x;
  const float* xu = x + base;
  // row means (thread per row); row 0 first so every thread can centre it
  float m = 0.f;
  if (tid < Rp) {
    const float4* row = (const float4*)(xu + (long)tid * D);
#pragma unroll 8
    for (int q = 0; q < D / 4; ++q) {
      const float4 v = row[q];
      m += (v.x + v.y) + (v.z + v.w);
    }
    m *= 1.f / D;
    mu[tid] = m;
  }
  __syncthreads();
  for (int f = tid; f < D; f += 256) d0s[f] = xu[f] - mu[0];
  __syncthreads();
  // <d_r, d_0> and |d_r|^2 per row
  float dot = 0.f, sq = 0.f;
  if (tid < Rp) {
    const float4* row = (const float4*)(xu + (long)tid * D);
#pragma unroll 8
    for (int q = 0; q < D / 4; ++q) {
      const float4 v = row[q];
      const float a0 = v.x - m, a1 = v.y - m, a2 = v.z - m, a3 = v.w - m;
      dot += a0 * d0s[4 * q] + a1 * d0s[4 * q + 1] + a2 * d0s[4 * q + 2] + a3 * d0s[4 * q + 3];
      sq += a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3;
    }
  }
  // unit scalars: v, lap_v, J_v (per row), s, s', s''
  const bool jrow = tid >= ROW_J0 && tid < R;
  float p_lap = jrow ? sq : 0.f;  // sum_c |J_d|^2
  if (tid == ROW_LAP) p_lap += dot;
  p_lap = wave_sum(p_lap);
  if ((tid & 63) == 0) red[tid >> 6] = p_lap;
  if (tid == 0) red[4] = dot;  // <d0, d0>
  __syncthreads();
  const float v = red[4] / D;
  const float lapv = 2.f / D * (red[0] + red[1] + red[2] + red[3]);
  const float ve = v + eps;
  const float sv = rsqrtf(ve);
  const float s1 = -0.5f * sv / ve;
  const float s2 = 0.75f * sv / (ve * ve);
  const float jv = jrow ? 2.f / D * dot : 0.f;
  float p_jv2 = wave_sum(jv * jv);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = p_jv2;
  js[tid] = s1 * jv;
  __syncthreads();
  const float laps = s1 * lapv + s2 * (red[0] + red[1] + red[2] + red[3]);
  // output: thread per feature, rows streamed
  float* yu = y + base;
  for (int f = tid; f < D; f += 256) {
    const float g = scale[f], d0 = d0s[f];
    float cross = 0.f;
#pragma unroll 8
    for (int r = ROW_J0; r < R; ++r) {
      const float dr = xu[(long)r * D + f] - mu[r];
      cross += dr * js[r];
      yu[(long)r * D + f] = g * (dr * sv + d0 * js[r]);
    }
    for (int r = R > ROW_J0 ? R : ROW_J0; r < Rp; ++r) yu[(long)r * D + f] = 0.f;
    yu[f] = g * d0 * sv + bias[f];
    if (Rp > 1) {
      const float dl = xu[D + f] - mu[1];
      yu[D + f] = g * (dl * sv + d0 * laps + 2.f * cross);
    }
  }
}

// Value rows only (Rp = 1, the Metropolis forward pass): y = scale (x - mu)
// s + bias with s = 1 / sqrt(mean((x - mu)^2) + eps) -- the rule above without
// tangent rows.  One wavefront per row (D / 64 features per lane), four rows
// per workgroup; mean and variance are wavefront sums.
template <int D>
__global__ __launch_bounds__(256) void layernorm_fwd_kernel(long rows, const float* __restrict__ x,
                                                            float* __restrict__ y, const float* __restrict__ scale,
                                                            const float* __restrict__ bias, float eps) {
  constexpr int PL = D / 64;  // features per lane
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  const int f0 = lane * PL;
  float v[PL];
#pragma unroll
  for (int e = 0; e < PL; ++e) v[e] = x[row * D + f0 + e];
  float m = 0.f;
#pragma unroll
  for (int e = 0; e < PL; ++e) m += v[e];
  m = wave_sum(m) * (1.f / D);
  float sq = 0.f;
#pragma unroll
  for (int e = 0; e < PL; ++e) {
    v[e] -= m;
    sq += v[e] * v[e];
  }
  const float sv = rsqrtf(wave_sum(sq) * (1.f / D) + eps);
#pragma unroll
  for (int e = 0; e < PL; ++e) y[row * D + f0 + e] = scale[f0 + e] * (v[e] * sv) + bias[f0 + e];
}

// The same rule with the unit's Rp x D tile read from HBM once into LDS, used
// when the tile fits.  Rows are D + 4 floats apart, so every LDS access is a
// 16-byte access: the row-per-thread statistics passes read four features at
// a time (a row's start moves by one 16-byte slot per thread: no bank
// conflicts), and the output pass gives each thread four consecutive features
// of every (NTH / (D/4))-th row, written back as 16-byte stores.  NTH = 512
// when the tile leaves room for one workgroup per CU only (two waves per
// SIMD instead of one), else 256.  Rp <= 256 (one thread per row in the
// statistics passes).
template <int D, int NTH>
__global__ __launch_bounds__(NTH) void layernorm_lap_lds_kernel(int Rp, int R, const float* __restrict__ x,
                                                                float* __restrict__ y, const float* __restrict__ scale,
                                                                const float* __restrict__ bias, float eps) {
  constexpr int SD = D + 4;
  constexpr int F4 = D / 4;       // float4 per row
  constexpr int RG = NTH / F4;    // row groups of the output pass
  constexpr int NW = NTH / 64;    // waves
  static_assert(NTH % F4 == 0 && NTH >= 256, "D must divide 4 * NTH");
  extern __shared__ __attribute__((aligned(16))) float tile[];  // Rp x SD
  __shared__ __attribute__((aligned(16))) float d0s[D];
  __shared__ float mu[NTH], js[NTH], red[NW + 1];
  __shared__ __attribute__((aligned(16))) float crs[RG * D];  // partial cross sums per row group
  const long base = (long)blockIdx.x * Rp * D;
  const int tid = threadIdx.x;
  const float4* x4 = (const float4*)(x + base);
#pragma unroll 8
  for (int idx = tid; idx < Rp * F4; idx += NTH) {
    const int r = idx / F4, q = idx - r * F4;
    *(float4*)(tile + r * SD + 4 * q) = x4[idx];
  }
  __syncthreads();
  float m = 0.f;
  if (tid < Rp) {
    const float4* row = (const float4*)(tile + tid * SD);
    float m4[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
    for (int q = 0; q < F4; ++q) {
      const float4 v = row[q];
      m4[0] += v.x;
      m4[1] += v.y;
      m4[2] += v.z;
      m4[3] += v.w;
    }
    m = ((m4[0] + m4[1]) + (m4[2] + m4[3])) * (1.f / D);
    mu[tid] = m;
  }
  __syncthreads();
  for (int f = tid; f < D; f += NTH) d0s[f] = tile[f] - mu[0];
  __syncthreads();
  float dot = 0.f, sq = 0.f;
  if (tid < Rp) {
    const float4* row = (const float4*)(tile + tid * SD);
    const float4* d04 = (const float4*)d0s;
    float dt[4] = {0.f, 0.f, 0.f, 0.f}, s4[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
    for (int q = 0; q < F4; ++q) {
      const float4 v = row[q], d = d04[q];
      const float a0 = v.x - m, a1 = v.y - m, a2 = v.z - m, a3 = v.w - m;
      dt[0] += a0 * d.x;
      dt[1] += a1 * d.y;
      dt[2] += a2 * d.z;
      dt[3] += a3 * d.w;
      s4[0] += a0 * a0;
      s4[1] += a1 * a1;
      s4[2] += a2 * a2;
      s4[3] += a3 * a3;
    }
    dot = (dt[0] + dt[1]) + (dt[2] + dt[3]);
    sq = (s4[0] + s4[1]) + (s4[2] + s4[3]);
  }
  const bool jrow = tid >= ROW_J0 && tid < R;
  float p_lap = jrow ? sq : 0.f;
  if (tid == ROW_LAP) p_lap += dot;
  p_lap = wave_sum(p_lap);
  if ((tid & 63) == 0) red[tid >> 6] = p_lap;
  if (tid == 0) red[NW] = dot;
  __syncthreads();
  const float v = red[NW] / D;
  float lap_sum = red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) lap_sum += red[w];
  const float lapv = 2.f / D * lap_sum;
  const float ve = v + eps;
  const float sv = rsqrtf(ve);
  const float s1 = -0.5f * sv / ve;
  const float s2 = 0.75f * sv / (ve * ve);
  const float jv = jrow ? 2.f / D * dot : 0.f;
  float p_jv2 = wave_sum(jv * jv);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = p_jv2;
  js[tid] = s1 * jv;
  __syncthreads();
  float jv2_sum = red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) jv2_sum += red[w];
  const float laps = s1 * lapv + s2 * jv2_sum;
  float* yu = y + base;
  const int f4 = tid % F4, ro = tid / F4, f = 4 * f4;
  const float4 dd = *(const float4*)&d0s[f];
  const float g[4] = {scale[f], scale[f + 1], scale[f + 2], scale[f + 3]}, d0[4] = {dd.x, dd.y, dd.z, dd.w};
  float cr[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
  for (int r = ROW_J0 + ro; r < R; r += RG) {
    const float4 t = *(const float4*)(tile + r * SD + f);
    const float mr = mu[r], jr = js[r];
    const float dr[4] = {t.x - mr, t.y - mr, t.z - mr, t.w - mr};
    float o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cr[q] += dr[q] * jr;
      o[q] = g[q] * (dr[q] * sv + d0[q] * jr);
    }
    *(float4*)(yu + (long)r * D + f) = float4{o[0], o[1], o[2], o[3]};
  }
  for (int r = (R > ROW_J0 ? R : ROW_J0) + ro; r < Rp; r += RG)
    *(float4*)(yu + (long)r * D + f) = float4{0.f, 0.f, 0.f, 0.f};
  *(float4*)&crs[ro * D + f] = float4{cr[0], cr[1], cr[2], cr[3]};
  __syncthreads();
  if (ro == 0) {
    const float bb[4] = {bias[f], bias[f + 1], bias[f + 2], bias[f + 3]};
    float ov[4], ol[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float cross = 0.f;
#pragma unroll
      for (int gq = 0; gq < RG; ++gq) cross += crs[gq * D + f + q];
      ov[q] = g[q] * d0[q] * sv + bb[q];
      const float dl = tile[SD + f + q] - mu[1];
      ol[q] = g[q] * (dl * sv + d0[q] * laps + 2.f * cross);
    }
    *(float4*)(yu + f) = float4{ov[0], ov[1], ov[2], ov[3]};
    if (Rp > 1) *(float4*)(yu + D + f) = float4{ol[0], ol[1], ol[2], ol[3]};
  }
}

// Softmax forward Laplacian over keys j for query i of (walker, head) z:
//   s_j: value S1[(i,0)][j]; J_c = S1[(i,c)][j] + S2[(j,c)][i];
//        lap = S1[(i,1)][j] + S2[(j,1)][i] + 2 S3[i][j]
//   a = softmax(s);  mu_c = sum_j a_j J_jc;  delta_jc = J_jc - mu_c
//   J_a_jc = a_j delta_jc
//   lap_a_j = a_j (sum_c delta_jc^2 - sum_k a_k sum_c delta_kc^2 + lap_j - sum_k a_k lap_k)
//
// One workgroup per (z, i); N <= 64 keys, one per lane.  Wave w takes the
// Jacobian rows c = 2 + w, 6 + w, ...: lane j gathers J_jc, mu_c is a wave
// sum, the centred delta_jc goes to LDS for the output pass and its square
// is accumulated per lane; the four waves' partial sums give sum_c delta_jc^2.
__global__ __launch_bounds__(256) void softmax_lap_kernel(SoftmaxArgs a) {
  extern __shared__ float sm[];
  const int N = a.N, Rp = a.Rp, R = a.R;
  float* J = sm;              // Rp x N: delta_jc (rows c >= 2 used)
  float* av = J + Rp * N;     // N: a_j
  float* lp = av + N;         // N: lap_j
  float* rq = lp + N;         // 4 x 64: per-wave partial sum_c delta_jc^2, then rq[j] = total
  float* red = rq + 256;      // 8
  // Workgroups go to the eight XCDs round-robin; the N query rows of one
  // (walker, head) are given to workgroups 8 apart so that they share one
  // XCD's L2 for their strided reads of S2[z].
  const int b = blockIdx.x, x8 = b & 7, slot = b >> 3;
  const int z = (slot / N) * 8 + x8, i = slot % N;
  if (z >= a.nz) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long zS = (long)z * N * Rp * a.lds;
  const float* S1 = a.S1 + zS + (long)i * Rp * a.lds;
  const float* S2 = a.S2 + zS + i;
  for (int j = tid; j < N; j += 256) {
    av[j] = S1[j];
    if (Rp > 1) lp[j] = S1[a.lds + j] + S2[((long)j * Rp + 1) * a.lds] + 2.f * a.S3[(long)z * N * a.lds + (long)i * a.lds + j];
  }
  __syncthreads();
  if (tid < 64) {
    float m = -INFINITY;
    for (int j = tid; j < N; j += 64) m = fmaxf(m, av[j]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    float zsum = 0.f;
    for (int j = tid; j < N; j += 64) zsum += expf(av[j] - m);
    zsum = wave_sum(zsum);
    for (int j = tid; j < N; j += 64) av[j] = expf(av[j] - m) / zsum;
  }
  __syncthreads();
  float* A1 = a.A1 + (long)z * N * Rp * a.lda1 + (long)i * Rp * a.lda1;
  if (Rp == 1) {
    for (int j = tid; j < a.lda1; j += 256) A1[j] = j < N ? av[j] : 0.f;
    return;
  }
  // Jacobian rows: centring by mu_c and the per-key sum of squares
  const bool jok = lane < N;
  const float aj = jok ? av[lane] : 0.f;
  const int jl = jok ? lane : 0;  // clamped key index: every address stays in the buffers
  const float* S2j = S2 + (long)jl * Rp * a.lds;
  float rpart = 0.f;
#pragma unroll 4
  for (int c = ROW_J0 + wave; c < R; c += 4) {
    const float jv = jok ? S1[(long)c * a.lds + jl] + S2j[(long)c * a.lds] : 0.f;
    const float mu = wave_sum(aj * jv);
    const float dj = jok ? jv - mu : 0.f;
    if (jok) J[c * N + lane] = dj;
    rpart += dj * dj;
  }
  rq[wave * 64 + lane] = rpart;
  __syncthreads();
  if (tid < 64) {
    const float r = rq[lane] + rq[64 + lane] + rq[128 + lane] + rq[192 + lane];
    const float nu = wave_sum(jok ? av[lane] * r : 0.f);
    const float lam = wave_sum(jok ? av[lane] * lp[lane] : 0.f);
    rq[lane] = r;
    if (tid == 0) {
      red[0] = nu;
      red[1] = lam;
    }
  }
  __syncthreads();
  const float nu = red[0], lam = red[1];
  // A1[c][j]: wave w writes rows c = w, w + 4, ...; lanes walk j (no index division)
  for (int c = wave; c < Rp; c += 4) {
    float* row = A1 + (long)c * a.lda1;
    for (int j = lane; j < a.lda1; j += 64) {
      float v = 0.f;
      if (j < N) {
        if (c == ROW_VAL)
          v = av[j];
        else if (c == ROW_LAP)
          v = av[j] * (rq[j] - nu + lp[j] - lam);
        else if (c < R)
          v = av[j] * J[c * N + j];
      }
      row[j] = v;
    }
  }
  if (a.A2) {
    // A2[(j, c)] = 2 a_j delta_jc, key-major; zero past N * Rp up to the 32-padding
    const long lda2 = ((long)N * Rp + 31) / 32 * 32;
    float* A2 = a.A2 + (long)z * N * lda2 + (long)i * lda2;
    for (int j = wave; j < N; j += 4) {
      const float two_a = 2.f * av[j];
      for (int c = lane; c < Rp; c += 64) A2[(long)j * Rp + c] = (c >= ROW_J0 && c < R) ? two_a * J[c * N + j] : 0.f;
    }
    for (long idx = (long)N * Rp + tid; idx < lda2; idx += 256) A2[idx] = 0.f;
  }
}

}  // namespace

int launch_psi_features(const FermiDims& d, const float* elec, const float* atoms, float* X, int kin, int rescale,
                        hipStream_t s) {
  const long total = (long)d.Bc * d.N * d.Rp * kin;
  hipLaunchKernelGGL(psi_features_kernel, grid_for(total), 256, 0, s, d, elec, atoms, X, kin, rescale);
  return 0;
}

// static LDS of layernorm_lap_lds_kernel<D, NTH> in bytes (d0s, mu, js, red,
// crs), with room for alignment
template <int D, int NTH>
constexpr size_t ln_lds_fixed() {
  return (size_t)(D + 2 * NTH + NTH / 64 + 1 + (NTH / (D / 4)) * D) * sizeof(float) + 64;
}

template <int D, int NTH>
bool launch_ln_lds_n(int units, int Rp, int R, const float* x, float* y, const float* scale, const float* bias,
                     float eps, hipStream_t s) {
  const size_t smem = (size_t)Rp * (D + 4) * sizeof(float);
  const size_t fixed = ln_lds_fixed<D, NTH>();
  if (Rp > 256 || smem + fixed > 160 * 1024) return false;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)layernorm_lap_lds_kernel<D, NTH>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)(160 * 1024 - fixed));
    attr = true;
  }
  hipLaunchKernelGGL((layernorm_lap_lds_kernel<D, NTH>), dim3(units), dim3(NTH), smem, s, Rp, R, x, y, scale, bias,
                     eps);
  return true;
}

// 512 threads when the tile allows one workgroup per CU only, else 256
template <int D>
bool launch_ln_lds(int units, int Rp, int R, const float* x, float* y, const float* scale, const float* bias,
                   float eps, hipStream_t s) {
  const size_t smem = (size_t)Rp * (D + 4) * sizeof(float);
  if (smem + ln_lds_fixed<D, 256>() > 80 * 1024 &&
      launch_ln_lds_n<D, 512>(units, Rp, R, x, y, scale, bias, eps, s))
    return true;
  return launch_ln_lds_n<D, 256>(units, Rp, R, x, y, scale, bias, eps, s);
}

int launch_layernorm_lap(int units, int Rp, int R, int D, const float* x, float* y, const float* scale,
                         const float* bias, float eps, hipStream_t s) {
  if (Rp > 256) return -1;
  if (Rp == 1 && units > 0) {  // value rows only
    const unsigned g = (unsigned)((units + 3) / 4);
    switch (D) {
      case 64: hipLaunchKernelGGL(layernorm_fwd_kernel<64>, dim3(g), dim3(256), 0, s, (long)units, x, y, scale, bias, eps); return 0;
      case 128: hipLaunchKernelGGL(layernorm_fwd_kernel<128>, dim3(g), dim3(256), 0, s, (long)units, x, y, scale, bias, eps); return 0;
      case 256: hipLaunchKernelGGL(layernorm_fwd_kernel<256>, dim3(g), dim3(256), 0, s, (long)units, x, y, scale, bias, eps); return 0;
      case 512: hipLaunchKernelGGL(layernorm_fwd_kernel<512>, dim3(g), dim3(256), 0, s, (long)units, x, y, scale, bias, eps); return 0;
    }
    return -1;
  }
  if (Rp > 1 && units > 0) {  // Laplacian mode: tile resident in LDS when it fits
    bool ok = false;
    if (D == 64) ok = launch_ln_lds<64>(units, Rp, R, x, y, scale, bias, eps, s);
    if (D == 128) ok = launch_ln_lds<128>(units, Rp, R, x, y, scale, bias, eps, s);
    if (D == 256) ok = launch_ln_lds<256>(units, Rp, R, x, y, scale, bias, eps, s);
    if (ok) return 0;
  }
  switch (D) {
    case 64: hipLaunchKernelGGL(layernorm_lap_kernel<64>, dim3(units), dim3(256), 0, s, Rp, R, x, y, scale, bias, eps); return 0;
    case 128: hipLaunchKernelGGL(layernorm_lap_kernel<128>, dim3(units), dim3(256), 0, s, Rp, R, x, y, scale, bias, eps); return 0;
    case 256: hipLaunchKernelGGL(layernorm_lap_kernel<256>, dim3(units), dim3(256), 0, s, Rp, R, x, y, scale, bias, eps); return 0;
    case 512: hipLaunchKernelGGL(layernorm_lap_kernel<512>, dim3(units), dim3(256), 0, s, Rp, R, x, y, scale, bias, eps); return 0;
  }
  return -1;
}

int launch_softmax_lap(const SoftmaxArgs& a, hipStream_t s) {
  const size_t smem = ((size_t)a.Rp * a.N + 2 * a.N + 256 + 8) * 4;
  if (smem > 160 * 1024 || a.N > 64) return -1;
  hipLaunchKernelGGL(softmax_lap_kernel, dim3((unsigned)(round_up(a.nz, 8) * a.N)), dim3(256), smem, s, a);
  return 0;
}
