// FP64 determinants of the orbital matrices and their forward Laplacian.
//
// The orbital matrix A[i][o] = phi[i][o] * env[i][o] (electron i, orbital o)
// is formed in FP64 from the FP32 orbital-projection output and the envelope
// (product rule, laplacian/primitives/core.py:494-500); log|det A| and its
// sign are those of jnp.linalg.slogdet (wavefunction/output/logdet.py:65).
//
// Laplacian mode restates laplacian/primitives/slogdet.py:44-59:
//   d logdet / dx_c = tr(A^-1 J_c)
//   lap logdet      = tr(A^-1 L) - sum_c tr((A^-1 J_c)^2)
// with B_c = A^-1 J_c on FP64 MFMA.  Three kernels: det_small_kernel
// (N <= 16, one wavefront per determinant), det_mfma_kernel (16 < N <= 64,
// Laplacian mode, one workgroup per determinant) and det_fwd_reg_kernel
// (16 < N <= 64, log|det| and sign only, one wavefront per determinant with
// the matrix in registers).

#include <algorithm>
#include <type_traits>

#include "kernels.h"

namespace {

// log|prod_p v_p| over the pivots without one log per factor: the running
// product is kept as a mantissa in [0.5, 1) and a binary exponent (both
// factors split with frexp, so it neither overflows nor underflows), and one
// log is taken at the end.  A zero pivot gives -inf, a NaN or inf pivot
// propagates, as the sum of logs would.
struct LogAbsProduct {
  double m = 1.0;
  int e = 0;
  __device__ void mul(double v) {
    int ev, em;
    const double mv = frexp(fabs(v), &ev);
    m = frexp(m * mv, &em);
    e += ev + em;
  }
  __device__ double log_abs() const { return log(m) + (double)e * 0.69314718055994530942; }
};

// Orbital-matrix element (value, and optionally Jacobian row c or Laplacian).
__device__ inline const float* env_ptr(const FermiDims& d, const float* env, int w, int i, int k, int o) {
  return env + ((((long)w * d.N + i) * d.K + k) * d.N + o) * env_stride(d);
}

// ---------------------------------------------------------------------------
// Larger systems (16 < N <= 64): the Laplacian traces with FP64 MFMA.
// Gauss-Jordan with the augmented matrix in registers (see below).  The coordinate loop
// then runs without workgroup barriers: wave q owns coordinates c = q, q+4, ...
// and forms every 16x16 tile of B_c = A^-1 J_c in its accumulators with
// v_mfma_f64_16x16x4f64.  The A operand (rows of A^-1) comes from LDS; the B
// operand, J_c = dPhi/dx_c * env, comes straight from HBM into registers, one
// coordinate ahead of its use, and is multiplied in FP64 by the envelope
// values env[i][o] held in LDS.  The envelope-derivative part of J_c is nonzero
// only in row c/3 and enters as the rank-1 update B_c += A^-1[:, c/3] u^T.
// tr(B_c) is read from the diagonal tiles and tr(B_c^2) = sum_{PQ} <B_PQ,
// B_QP^T> from per-wave transposes in LDS, one column block of tiles at a time.
// MFMA operand layout (16x16x4, f64): lane (r = lane%16, g = lane/16) supplies
// A[r][4kk + g] and B[4kk + g][r]; result element v of the lane is row g + 4v,
// column r (this differs from the f32 16x16x4 result, rows 4g + v; checked on
// the device element by element).
// ---------------------------------------------------------------------------
template <int NP>
constexpr int det_mfma_inv_stride() {
  return NP + 2;  // rows 2 doubles apart mod 32: conflict-free A-operand reads
}
constexpr int DM_TS = 17;  // transpose tile stride (doubles)
template <int NP>
constexpr int det_mfma_env_stride() {
  return NP % 64 == 0 ? NP + 16 : NP % 32 == 0 ? NP + 16 : NP;  // == 16 or 48 mod 64 floats: rows g hit disjoint banks
}
template <int NP>
constexpr size_t det_mfma_smem() {
  return (size_t)((NP * NP > 4 * (NP / 16) * 16 * DM_TS ? NP * NP : 4 * (NP / 16) * 16 * DM_TS) +
                  NP * det_mfma_inv_stride<NP>() + 8) * 8 +
         (size_t)NP * det_mfma_env_stride<NP>() * 4 + 16;
}

template <int NP>
__global__ __launch_bounds__(256, (NP <= 48 ? 2 : 1)) void det_mfma_kernel(FermiDims d, const float* __restrict__ Phi, long ldphi,
                                                       const float* __restrict__ env, double* __restrict__ out) {
  constexpr int IS = det_mfma_inv_stride<NP>();
  constexpr int NT = NP / 16, KK = NP / 4;
  constexpr int TWS = NT * 16 * DM_TS;  // a wave's transpose tiles: one column block of B_c
  constexpr int AREG = NP * NP > 4 * TWS ? NP * NP : 4 * TWS;
  extern __shared__ double sm[];
  double* A = sm;           // NP*NP for the elimination, then the wave transpose tiles
  double* Inv = A + AREG;   // NP*IS
  double* red = Inv + NP * IS;
  constexpr int ES = det_mfma_env_stride<NP>();
  float* E0 = (float*)(red + 8);  // NP*ES envelope values env[i][o]
  int* ipiv = (int*)(E0 + NP * ES);
  const int N = d.N, T = 3 * N;
  // the K determinants of a walker share rows of Phi: same XCD (L2)
  int w, k;
  xcd_tile(blockIdx.x, d.K, &w, &k);
  if (w >= d.Bc) return;
  const long det_id = (long)w * d.K + k;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long rowstride = (long)d.Rp * ldphi;  // floats per unit
  const long kN = (long)k * N;

#pragma unroll 4
  for (int e = tid; e < NP * NP; e += 256) {
    const int i = e / NP, o = e % NP;
    double a = (i == o) ? 1.0 : 0.0;
    float e0 = 0.f;
    if (i < N && o < N) {
      e0 = env_ptr(d, env, w, i, k, o)[0];
      a = (double)Phi[(long)unit_idx(d, w, i) * rowstride + kN + o] * (double)e0;
    }
    A[e] = a;
    E0[i * ES + o] = e0;
  }
  __syncthreads();

  // ---- Gauss-Jordan on [A | I] with the matrix in registers ----
  // Thread (ti, tj) = (tid % 16, tid / 16) owns rows i = ti + 16a and columns
  // j = tj + 16b of the NP x 2NP augmented matrix.  No rows are exchanged:
  // the pivot for column p is the unused row q with the largest |A[q][p]|
  // (smallest q on ties), and at the end row p of A^-1 is row q of the right
  // half.  Per pivot: the owners of column p (16 lanes of one DPP row) have
  // already reduced their candidates to (q, A[q][p]); barrier; the owners of
  // row q and of column p publish them; barrier; every thread updates its
  // elements and the owners of column p + 1 reduce the next candidates.
  constexpr int RA = NP / 16, CB = NP / 8;
  const int ti = tid & 15, tj = tid >> 4;
  double x[RA][CB];
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int b = 0; b < CB; ++b) {
      const int i = ti + 16 * a, j = tj + 16 * b;
      x[a][b] = j < NP ? A[i * NP + j] : (i == j - NP ? 1.0 : 0.0);
    }
  __syncthreads();  // A is scratch from here on
  double* prow = A;             // pivot row, 2NP
  double* pcol = A + 2 * NP;    // pivot column, NP
  double* cval = A + 3 * NP;    // signed pivot value
  int* cidx = (int*)(A + 3 * NP + 1);
  int* perm = (int*)(A + 3 * NP + 2);  // perm[p] = pivot row of column p
  unsigned used = 0;
  int pos[RA];  // pivot index of the thread's rows
  // candidates of a column held in x[.][b] by the 16 lanes of this DPP row
  auto candidates = [&](int b) __attribute__((always_inline)) {
    double v = -1.0, sv = 0.0;
    int bi = NP;
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      const double y = fabs(x[a][b]);
      if (!((used >> a) & 1u) && (bi == NP || y > v)) {  // any unused row, even a non-finite one
        v = y;
        sv = x[a][b];
        bi = ti + 16 * a;
      }
    }
    // lanes without an unused row (bi == NP) take any offered candidate, so
    // the result is a valid unused row whenever one exists
    auto pick = [&](double ov, double osv, int oi) __attribute__((always_inline)) {
      if (oi != NP && (bi == NP || ov > v || (ov == v && oi < bi))) {
        v = ov;
        sv = osv;
        bi = oi;
      }
    };
    pick(dpp_d<0xB1>(v), dpp_d<0xB1>(sv), __builtin_amdgcn_mov_dpp(bi, 0xB1, 0xF, 0xF, false));
    pick(dpp_d<0x4E>(v), dpp_d<0x4E>(sv), __builtin_amdgcn_mov_dpp(bi, 0x4E, 0xF, 0xF, false));
    pick(dpp_d<0x141>(v), dpp_d<0x141>(sv), __builtin_amdgcn_mov_dpp(bi, 0x141, 0xF, 0xF, false));
    pick(dpp_d<0x140>(v), dpp_d<0x140>(sv), __builtin_amdgcn_mov_dpp(bi, 0x140, 0xF, 0xF, false));
    if (ti == 0) {
      *cval = sv;
      *cidx = bi;
    }
  };
  if (tj == 0) candidates(0);
  LogAbsProduct logdet;
  int sgn = 1;
#pragma unroll
  for (int pb = 0; pb < RA; ++pb) {
    for (int pr = 0; pr < 16; ++pr) {
      const int p = 16 * pb + pr;
      __syncthreads();
      const int q = *cidx;
      const double piv = *cval;
      const double rp = 1.0 / piv;
      if (tid == 0) {
        logdet.mul(piv);
        if (piv < 0) sgn = -sgn;
        perm[p] = q;
      }
      if (ti == (q & 15)) {
#pragma unroll
        for (int a = 0; a < RA; ++a)
          if (a == (q >> 4)) {
#pragma unroll
            for (int b = 0; b < CB; ++b) prow[tj + 16 * b] = x[a][b];
          }
      }
      if (tj == pr) {
#pragma unroll
        for (int a = 0; a < RA; ++a) pcol[ti + 16 * a] = x[a][pb];
      }
      __syncthreads();
      double r[CB];
#pragma unroll
      for (int b = 0; b < CB; ++b) r[b] = prow[tj + 16 * b] * rp;
#pragma unroll
      for (int a = 0; a < RA; ++a) {
        const int i = ti + 16 * a;
        if (i == q) {
#pragma unroll
          for (int b = 0; b < CB; ++b) x[a][b] = r[b];
          used |= 1u << a;
          pos[a] = p;
        } else {
          const double f = pcol[i];
#pragma unroll
          for (int b = 0; b < CB; ++b) x[a][b] -= f * r[b];
        }
      }
      if (p + 1 < NP && tj == ((p + 1) & 15)) {
        if (pr == 15) {
          candidates(pb + 1 < RA ? pb + 1 : pb);  // p + 1 < NP implies pb + 1 < RA here
        } else {
          candidates(pb);
        }
      }
    }
  }
  // row p of A^-1 = right half of row perm[p]
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int b = 0; b < CB; ++b) {
      const int j = tj + 16 * b;
      if (j >= NP) Inv[pos[a] * IS + (j - NP)] = x[a][b];
    }
  __syncthreads();
  if (tid == 0) {
    // sign of the row permutation: parity from its cycles
    unsigned long long seen = 0ull;
    for (int p = 0; p < NP; ++p) {
      if ((seen >> p) & 1ull) continue;
      int len = 0, c = p;
      while (len < NP && c < NP && !((seen >> c) & 1ull)) {
        seen |= 1ull << c;
        c = perm[c];
        ++len;
      }
      if ((len & 1) == 0) sgn = -sgn;
    }
  }
  double* o_ = out + det_id * (3 + T);

  // tr(A^-1 L), L = lap(phi env) = lap(phi) env + phi lap(env) + 2 grad(phi).grad(env)
  double part = 0.0;
#pragma unroll 4
  for (int e = tid; e < N * N; e += 256) {
    const int i = e / N, o = e % N;
    const float* ph = Phi + (long)unit_idx(d, w, i) * rowstride + kN + o;
    const float* ep = env_ptr(d, env, w, i, k, o);
    double L = (double)ph[ROW_LAP * ldphi] * ep[0] + (double)ph[0] * ep[1];
    for (int ax = 0; ax < 3; ++ax) L += 2.0 * (double)ph[(ROW_J0 + 3 * i + ax) * ldphi] * ep[2 + ax];
    part += Inv[o * IS + i] * L;
  }
  part = wave_sum_d(part);
  if (lane == 0) red[wave] = part;
  __syncthreads();  // also: A is free for the transpose tiles from here on
  const double trAL = red[0] + red[1] + red[2] + red[3];
  __syncthreads();

  // ---- coordinate loop, no workgroup barriers ----
  const int lr = lane & 15, lg = lane >> 4;
  double* tile = A + wave * TWS;
  int urow[KK];  // unit * Rp + ROW_J0 for the lane's matrix row m = 4kk + lg, or -1 past N
  float jv[KK][NT];
#pragma unroll
  for (int kk = 0; kk < KK; ++kk) {
    const int m = 4 * kk + lg;
    urow[kk] = m < N ? unit_idx(d, w, m) * d.Rp + ROW_J0 : -1;
  }
  const float* Phik = Phi + kN;
  auto load_j = [&](int kk, int c) {
#pragma unroll
    for (int Q = 0; Q < NT; ++Q) {
      const int o = 16 * Q + lr;
      jv[kk][Q] = (c < T && urow[kk] >= 0 && o < N) ? Phik[(long)(urow[kk] + c) * ldphi + o] : 0.f;
    }
  };
#pragma unroll
  for (int kk = 0; kk < KK; ++kk) load_j(kk, wave);

  double tsum = 0.0;
  for (int c = wave; c < T; c += 4) {
    const int cn = c + 4;
    // envelope-derivative row: u[o] = phi[m*][o] d env[m*][o] / dx_c, m* = c / 3
    const int ms = c / 3, ax = c % 3;
    double u[NT];
    {
      const long ub = (long)unit_idx(d, w, ms) * rowstride + kN;
#pragma unroll
      for (int Q = 0; Q < NT; ++Q) {
        const int o = 16 * Q + lr;
        u[Q] = o < N ? (double)Phi[ub + o] * (double)env_ptr(d, env, w, ms, k, o)[2 + ax] : 0.0;
      }
    }
    f64x4 acc[NT][NT];
#pragma unroll
    for (int P = 0; P < NT; ++P)
#pragma unroll
      for (int Q = 0; Q < NT; ++Q) acc[P][Q] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      double av[NT], bv[NT];
#pragma unroll
      for (int P = 0; P < NT; ++P) av[P] = Inv[(16 * P + lr) * IS + 4 * kk + lg];
#pragma unroll
      for (int Q = 0; Q < NT; ++Q) bv[Q] = (double)jv[kk][Q] * (double)E0[(4 * kk + lg) * ES + 16 * Q + lr];
      load_j(kk, cn);
#pragma unroll
      for (int P = 0; P < NT; ++P)
#pragma unroll
        for (int Q = 0; Q < NT; ++Q) acc[P][Q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[P], bv[Q], acc[P][Q], 0, 0, 0);
    }
    // rank-1 envelope term and tr(B_c)
    double tr = 0.0;
#pragma unroll
    for (int P = 0; P < NT; ++P)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double g = Inv[(16 * P + lg + 4 * v) * IS + ms];
#pragma unroll
        for (int Q = 0; Q < NT; ++Q) acc[P][Q][v] += g * u[Q];
        if (lr == lg + 4 * v) tr += acc[P][P][v];
      }
    tr = wave_sum_d(tr);
    if (lane == 0) o_[3 + c] = tr;
    // tr(B_c^2) = sum_PQ <B_PQ, B_QP^T>: for each P the tiles (Q, P), Q >= P,
    // go through the wave's LDS slice together (one pair of wave barriers)
#pragma unroll
    for (int P = 0; P < NT; ++P) {
#pragma unroll
      for (int Q = P; Q < NT; ++Q)
#pragma unroll
        for (int v = 0; v < 4; ++v) tile[(Q - P) * 16 * DM_TS + (lg + 4 * v) * DM_TS + lr] = acc[Q][P][v];
      wave_sync();
#pragma unroll
      for (int Q = P; Q < NT; ++Q) {
        double s = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) s += acc[P][Q][v] * tile[(Q - P) * 16 * DM_TS + lr * DM_TS + lg + 4 * v];
        tsum += (P == Q) ? s : 2.0 * s;
      }
      wave_sync();
    }
  }
  tsum = wave_sum_d(tsum);
  if (lane == 0) red[4 + wave] = tsum;
  __syncthreads();
  if (tid == 0) {
    o_[0] = logdet.log_abs();
    o_[1] = (double)sgn;
    o_[2] = trAL - (red[4] + red[5] + red[6] + red[7]);
  }
}

// ---------------------------------------------------------------------------
// Forward mode, 16 < N <= 64: log|det| and sign with the matrix in registers.
// One wavefront per determinant, four per workgroup, no LDS: lane i holds row
// i of A (NP >= N doubles, NP a multiple of 8, zero-padded; lanes i >= N hold
// no row).  Elimination without row exchanges: the pivot of column p is the
// unused row with the largest |A[i][p]| (compared on the FP32 rounding of the
// magnitude with its low 6 bits replaced by the row, which picks the smaller
// row on ties and a row within 2^-17 of the largest otherwise; non-finite
// entries rank below every finite one, so a walker with a NaN still gets a
// valid pivot row), its row is read from that lane with readlane, and every
// other unused row subtracts its multiple of it from the columns after p.
// det = sign(sigma) * prod of pivots with sigma(p) = pivot row of column p; the
// parity of sigma is the number of unused rows below the pivot, summed over
// the steps.  The column loop is unrolled at compile time (column p is
// register p), so the update is NP - 1 - p fused multiply-adds per lane with
// the pivot-row entry in scalar registers.
// ---------------------------------------------------------------------------
__device__ inline double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Maximum of a 32-bit key over the wavefront: DPP within rows of 16, then the
// four row results through scalar registers.
__device__ inline unsigned wave_max_u32(unsigned k) {
  k = max(k, (unsigned)__builtin_amdgcn_mov_dpp((int)k, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  k = max(k, (unsigned)__builtin_amdgcn_mov_dpp((int)k, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  k = max(k, (unsigned)__builtin_amdgcn_mov_dpp((int)k, 0x141, 0xF, 0xF, false));  // row_half_mirror
  k = max(k, (unsigned)__builtin_amdgcn_mov_dpp((int)k, 0x140, 0xF, 0xF, false));  // row_mirror
  const unsigned r0 = (unsigned)__builtin_amdgcn_readlane((int)k, 0);
  const unsigned r1 = (unsigned)__builtin_amdgcn_readlane((int)k, 16);
  const unsigned r2 = (unsigned)__builtin_amdgcn_readlane((int)k, 32);
  const unsigned r3 = (unsigned)__builtin_amdgcn_readlane((int)k, 48);
  return max(max(r0, r1), max(r2, r3));
}

struct FwdLU {
  unsigned long long used;  // bit i: row i is a pivot row already (or holds no row)
  LogAbsProduct logdet;
  int neg, parity;
};

template <int P, int NP>
__device__ __forceinline__ void fwd_lu_steps(double (&m)[NP], int N, int lane, FwdLU& st) {
  if constexpr (P < NP) {
    if (P < N) {
      const double x = m[P];
      const double ax = fabs(x);
      const bool open = !((st.used >> lane) & 1ull);
      const unsigned tie = 63u - (unsigned)lane;
      const unsigned key =
          !open ? 0u : (ax <= 1.7976931348623157e308 ? ((__float_as_uint((float)ax) & ~63u) + 64u) | tie : tie);
      // key 0 decodes to lane 63: it wins only when every key is 0, i.e. the
      // one open row is a non-finite row 63
      const int piv = 63 - (int)(wave_max_u32(key) & 63u);
      const unsigned long long below = ~st.used & ((1ull << piv) - 1ull);
      st.parity ^= __builtin_popcountll(below) & 1;
      st.used |= 1ull << piv;
      const double pv = readlane_f64(x, piv);
      st.logdet.mul(pv);
      st.neg ^= pv < 0.0 ? 1 : 0;
      const double l = (open && lane != piv) ? x / pv : 0.0;
      // pivot-row entries read four at a time ahead of their multiply-adds, so
      // the scalar reads do not wait on each other
#pragma unroll
      for (int j0 = P + 1; j0 < NP; j0 += 4) {
        double r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (j0 + q < NP) r[q] = readlane_f64(m[j0 + q], piv);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (j0 + q < NP) m[j0 + q] = fma(-l, r[q], m[j0 + q]);
      }
    }
    fwd_lu_steps<P + 1, NP>(m, N, lane, st);
  }
}

template <int NP>
__global__ __launch_bounds__(256) void det_fwd_reg_kernel(FermiDims d, const float* __restrict__ Phi, long ldphi,
                                                          const float* __restrict__ env, double* __restrict__ out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int w, cb;
  xcd_tile(blockIdx.x, (d.K + 3) / 4, &w, &cb);
  const int k = 4 * cb + wv;
  if (w >= d.Bc || k >= d.K) return;  // the wave's own determinant
  const long det = (long)w * d.K + k;
  const int N = d.N;
  double m[NP];
  {
    const int i = lane < N ? lane : 0;
    const float* ph = Phi + (long)unit_idx(d, w, i) * d.Rp * ldphi + (long)k * N;
    const float* ep = env_ptr(d, env, w, i, k, 0);
#pragma unroll
    for (int j = 0; j < NP; ++j)
      m[j] = (lane < N && j < N) ? (double)ph[j] * (double)ep[j] : 0.0;  // value-only envelope (stride 1)
  }
  FwdLU st;
  st.used = N >= 64 ? 0ull : ~((1ull << N) - 1ull);
  st.neg = 0;
  st.parity = 0;
  fwd_lu_steps<0, NP>(m, N, lane, st);
  if (lane == 0) {
    out[det * 3] = st.logdet.log_abs();
    out[det * 3 + 1] = (st.neg ^ st.parity) ? -1.0 : 1.0;
    out[det * 3 + 2] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Small systems (N <= 16): one wavefront per determinant, four per workgroup,
// no workgroup barriers.  The augmented [A | I] (16 x 32 FP64) is reduced by
// Gauss-Jordan with partial pivoting in the wave's LDS slice; then, for each
// coordinate c, B_c = A^-1 J_c is one 16x16 FP64 MFMA tile (four K steps,
// A^-1 fragments in registers, J_c from HBM one coordinate pair ahead), with
// tr(B_c) from the diagonal lanes and tr(B_c^2) from a transposed read of the
// tile in the wave's LDS slice.
// ---------------------------------------------------------------------------

constexpr int SM_NP = 16;
// coordinates in flight in the small kernel (even).  Two, with the kernel
// bounded to 128 VGPRs (four waves per SIMD, 16 per CU), beat four at three
// waves per SIMD: more waves hide the LDS transposes and wave barriers.
constexpr int SM_PF = 2;
struct SmallDetSmem {
  double M[SM_NP * 2 * SM_NP];  // augmented matrix, later Inv in columns 16..31
  double Tt[2 * SM_NP * 17];    // transpose tiles of B_c, B_c+1 (odd row stride)
};
struct SmallFwdSmem {           // forward mode: the matrix and the pivot exchange area
  double M[SM_NP * 2 * SM_NP];
  double Tt[48];
};

template <bool LAPMODE>
__global__ __launch_bounds__(256, 4) void det_small_kernel(FermiDims d, const float* __restrict__ Phi, long ldphi,
                                                        const float* __restrict__ env, double* __restrict__ out) {
  constexpr bool lapmode = LAPMODE;
  using Smem = typename std::conditional<LAPMODE, SmallDetSmem, SmallFwdSmem>::type;
  __shared__ Smem smem[4];
  const int N = d.N, T = 3 * N;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // determinants 4cb .. 4cb + 3 of walker w; a walker's blocks share an XCD
  int w, cb;
  xcd_tile(blockIdx.x, (d.K + 3) / 4, &w, &cb);
  const int k = 4 * cb + wave;
  if (w >= d.Bc || k >= d.K) return;  // whole wave exits together
  const long det_id = (long)w * d.K + k;
  Smem& S = smem[wave];
  const int Rp = d.Rp;

  // per-lane elements e = lane + 64 t, t = 0..3: (i, o) = (e / 16, e % 16)
  float phi0[4], e0[4], ej[4][3];
  long base[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int e = lane + 64 * t, i = e >> 4, o = e & 15;
    const bool in = i < N && o < N;
    base[t] = -1;
    phi0[t] = 0.f;
    e0[t] = 0.f;
    ej[t][0] = ej[t][1] = ej[t][2] = 0.f;
    double a = (i == o) ? 1.0 : 0.0;
    if (in) {
      const int u = unit_idx(d, w, i);
      base[t] = (long)u * Rp * ldphi + k * N + o;
      const float* ep = env_ptr(d, env, w, i, k, o);
      phi0[t] = Phi[base[t]];
      e0[t] = ep[0];
      if (lapmode) {
        ej[t][0] = ep[2];
        ej[t][1] = ep[3];
        ej[t][2] = ep[4];
      }
      a = (double)phi0[t] * (double)e0[t];
    }
    S.M[i * 32 + o] = a;
    S.M[i * 32 + 16 + o] = (i == o) ? 1.0 : 0.0;
  }
  wave_sync();

  LogAbsProduct logdet;
  int sgn = 1;
  // pivot search over rows p..15 in lanes 0..15: DPP butterfly within the
  // first 16-lane row (largest |value|, smallest row on ties)
  auto pick = [&](double& v, int& bi, double ov, int oi) __attribute__((always_inline)) {
    if (oi != SM_NP && (bi == SM_NP || ov > v || (ov == v && oi < bi))) {
      v = ov;
      bi = oi;
    }
  };
  if constexpr (!lapmode) {
    // forward mode needs only log|det| and the sign: LU without row
    // exchanges with the 16x16 block in registers.  Lane (ti, tj) =
    // (lane % 16, lane / 16) owns row ti, columns tj + 4b (b = 0..3); the
    // pivot of column p is the unused row with the largest |entry| (DPP
    // inside the 16-lane row that holds column p); the pivot row and column go
    // through LDS and the rows not yet used are updated right of column p.
    // The parity of the row order enters the sign.
    const int ti = lane & 15, tj = lane >> 4;
    double x[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) x[b] = S.M[ti * 32 + tj + 4 * b];
    bool used = false;
    double* prow = S.Tt;       // 16 doubles
    double* pcol = S.Tt + 16;  // 16 doubles
    int* perm = (int*)(S.Tt + 32);
    for (int p = 0; p < SM_NP; ++p) {
      const int pb = p >> 2, pr = p & 3;
      double xp = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (b == pb) xp = x[b];
      double v = used ? -1.0 : fabs(xp);
      int bi = used ? SM_NP : ti;
      pick(v, bi, dpp_d<0xB1>(v), __builtin_amdgcn_mov_dpp(bi, 0xB1, 0xF, 0xF, false));
      pick(v, bi, dpp_d<0x4E>(v), __builtin_amdgcn_mov_dpp(bi, 0x4E, 0xF, 0xF, false));
      pick(v, bi, dpp_d<0x141>(v), __builtin_amdgcn_mov_dpp(bi, 0x141, 0xF, 0xF, false));
      pick(v, bi, dpp_d<0x140>(v), __builtin_amdgcn_mov_dpp(bi, 0x140, 0xF, 0xF, false));
      int q = __builtin_amdgcn_readlane(bi, 16 * pr);
      if (q >= SM_NP) q = 0;  // unreachable: an unused row always remains
      const double piv = readlane_d(xp, 16 * pr + q);
      logdet.mul(piv);
      if (piv < 0) sgn = -sgn;
      if (lane == 0) perm[p] = q;
      if (ti == q) {
#pragma unroll
        for (int b = 0; b < 4; ++b) prow[tj + 4 * b] = x[b];
      }
      if (tj == pr) pcol[ti] = xp;
      wave_sync();
      const double rp = 1.0 / piv;
      if (ti == q) {
        used = true;
      } else if (!used) {
        const double f = pcol[ti] * rp;
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if (tj + 4 * b > p) x[b] -= f * prow[tj + 4 * b];
      }
      wave_sync();
    }
    if (lane == 0) {
      int seen = 0;
      for (int p0 = 0; p0 < SM_NP; ++p0) {
        if ((seen >> p0) & 1) continue;
        int len = 0, c = p0;
        while (len < SM_NP && c < SM_NP && !((seen >> c) & 1)) {
          seen |= 1 << c;
          c = perm[c];
          ++len;
        }
        if ((len & 1) == 0) sgn = -sgn;
      }
      double* o_ = out + det_id * 3;
      o_[0] = logdet.log_abs();
      o_[1] = (double)sgn;
      o_[2] = 0.0;
    }
    return;
  }
  // Gauss-Jordan on [A | I] with the matrix in registers and no row
  // exchanges: lane (ti, tj) = (lane % 16, lane / 16) owns row ti, columns
  // tj + 4b (b = 0..7).  The pivot of column p is the unused row with the
  // largest |entry| (smallest row on ties), reduced with DPP inside the
  // 16-lane row that holds column p; the pivot row and column are published
  // through LDS; at the end row p of A^-1 is the right half of the pivot row
  // of column p, and the parity of that row permutation enters the sign.
  {
    const int ti = lane & 15, tj = lane >> 4;
    double x[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) x[b] = S.M[ti * 32 + tj + 4 * b];
    bool used = false;
    int pos = 0;
    double* prow = S.Tt;       // 32 doubles
    double* pcol = S.Tt + 32;  // 16 doubles
    int* perm = (int*)(S.Tt + 48);
    for (int p = 0; p < SM_NP; ++p) {
      const int pb = p >> 2, pr = p & 3;
      double xp = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (b == pb) xp = x[b];
      double v = used ? -1.0 : fabs(xp);
      int bi = used ? SM_NP : ti;
      pick(v, bi, dpp_d<0xB1>(v), __builtin_amdgcn_mov_dpp(bi, 0xB1, 0xF, 0xF, false));
      pick(v, bi, dpp_d<0x4E>(v), __builtin_amdgcn_mov_dpp(bi, 0x4E, 0xF, 0xF, false));
      pick(v, bi, dpp_d<0x141>(v), __builtin_amdgcn_mov_dpp(bi, 0x141, 0xF, 0xF, false));
      pick(v, bi, dpp_d<0x140>(v), __builtin_amdgcn_mov_dpp(bi, 0x140, 0xF, 0xF, false));
      int q = __builtin_amdgcn_readlane(bi, 16 * pr);
      if (q >= SM_NP) q = 0;  // unreachable: an unused row always remains
      const double piv = readlane_d(xp, 16 * pr + q);
      logdet.mul(piv);
      if (piv < 0) sgn = -sgn;
      if (lane == 0) perm[p] = q;
      if (ti == q) {
#pragma unroll
        for (int b = 0; b < 8; ++b) prow[tj + 4 * b] = x[b];
      }
      if (tj == pr) pcol[ti] = xp;
      wave_sync();
      const double rp = 1.0 / piv;
      const double f = ti == q ? 0.0 : pcol[ti];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const double r = prow[tj + 4 * b] * rp;
        x[b] = ti == q ? r : x[b] - f * r;
      }
      if (ti == q) {
        used = true;
        pos = p;
      }
      wave_sync();
    }
    // row pos of A^-1 = right half of this lane's row
#pragma unroll
    for (int b = 4; b < 8; ++b) S.M[pos * 32 + tj + 4 * b] = x[b];
    if (lane == 0) {
      int seen = 0;
      for (int p0 = 0; p0 < SM_NP; ++p0) {
        if ((seen >> p0) & 1) continue;
        int len = 0, c = p0;
        while (len < SM_NP && c < SM_NP && !((seen >> c) & 1)) {
          seen |= 1 << c;
          c = perm[c];
          ++len;
        }
        if ((len & 1) == 0) sgn = -sgn;
      }
    }
    wave_sync();
  }
  double* o_ = out + det_id * (3 + T);

  // tr(A^-1 L): lane's 4 elements (i, o) pair with Inv[o][i]
  double part = 0.0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (base[t] < 0) continue;
    const int e = lane + 64 * t, i = e >> 4, o = e & 15;
    const float* ep = env + ((((long)w * N + i) * d.K + k) * N + o) * 5;  // Laplacian mode: 5 floats
    double L = (double)Phi[base[t] + ROW_LAP * ldphi] * e0[t] + (double)phi0[t] * ep[1];
    for (int ax = 0; ax < 3; ++ax) L += 2.0 * (double)Phi[base[t] + (ROW_J0 + 3 * i + ax) * ldphi] * ej[t][ax];
    part += S.M[o * 32 + 16 + i] * L;
  }
  const double trAL = wave_sum_d(part);

  // Coordinate loop on v_mfma_f64_16x16x4f64: B_c = A^-1 J_c in four K
  // steps.  Lane (r, g) = (lane % 16, lane / 16) supplies A^-1[r][4kk + g]
  // (constant, in registers) and J_c[4kk + g][r] (from HBM, one coordinate
  // ahead); result element v is B_c[g + 4v][r].  The envelope-derivative part
  // of J_c lives in row m* = c / 3 only and enters as B_c += A^-1[:, m*] u^T.
  const int lr = lane & 15, lg = lane >> 4;
  // Coordinates c .. c + SM_PF - 1 are in flight from HBM: slot q of the
  // unrolled loop holds coordinate c0 + q and is refilled with c0 + q + SM_PF
  // once consumed (one coordinate of MFMA work is far shorter than an HBM
  // round trip).
  double av[4];
  long bb[4];
  float e0b[4], jv[SM_PF][4], uph[SM_PF], ude[SM_PF];
  // u for coordinate c: phi[m*][r] * d env[m*][r] / dx_c (lanes r < N)
  auto load_c = [&](int q, int c) __attribute__((always_inline)) {
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) jv[q][kk] = (c < T && bb[kk] >= 0) ? Phi[bb[kk] + (ROW_J0 + c) * ldphi] : 0.f;
    const int ms = c / 3;
    if (c < T && lr < N) {
      uph[q] = Phi[(long)unit_idx(d, w, ms) * Rp * ldphi + (long)k * N + lr];
      ude[q] = env[((((long)w * N + ms) * d.K + k) * N + lr) * 5 + 2 + c % 3];
    } else {
      uph[q] = 0.f;
      ude[q] = 0.f;
    }
  };
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int m = 4 * kk + lg;
    av[kk] = S.M[lr * 32 + 16 + m];
    const bool in = m < N && lr < N;
    bb[kk] = in ? (long)unit_idx(d, w, m) * Rp * ldphi + (long)k * N + lr : -1;
    e0b[kk] = in ? env[((((long)w * N + m) * d.K + k) * N + lr) * 5] : 0.f;
  }
#pragma unroll
  for (int q = 0; q < SM_PF; ++q) load_c(q, q);
  double* tile = S.Tt;
  double tsum = 0.0;
  wave_sync();
  // two coordinates per step: independent MFMA chains, reductions and
  // transposes interleave, and the step needs two wave barriers
  double* tileb = tile + 16 * 17;
  for (int c0 = 0; c0 < T; c0 += SM_PF)
#pragma unroll
  for (int q = 0; q < SM_PF; q += 2) {
    const int ca = c0 + q, cb = ca + 1;
    if (ca >= T) break;
    const bool hb = cb < T;
    double bva[4], bvb[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      bva[kk] = (double)jv[q][kk] * (double)e0b[kk];
      bvb[kk] = (double)jv[q + 1][kk] * (double)e0b[kk];  // zero when cb >= T
    }
    const double ua = (double)uph[q] * (double)ude[q], ub = (double)uph[q + 1] * (double)ude[q + 1];
    const int msa = ca / 3, msb = hb ? cb / 3 : 0;
    // refill both slots; coordinates past T load as zeros, so the second
    // half of a step past the end contributes nothing
    load_c(q, ca + SM_PF);
    load_c(q + 1, cb + SM_PF);
    f64x4 A = f64x4{0.0, 0.0, 0.0, 0.0}, B = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      A = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], bva[kk], A, 0, 0, 0);
      B = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], bvb[kk], B, 0, 0, 0);
    }
    double tra = 0.0, trb = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      A[v] += S.M[(lg + 4 * v) * 32 + 16 + msa] * ua;
      B[v] += S.M[(lg + 4 * v) * 32 + 16 + msb] * ub;
      if (lr == lg + 4 * v) {
        tra += A[v];
        trb += B[v];
      }
    }
    tra = wave_sum_dpp_d(tra);
    trb = wave_sum_dpp_d(trb);
    if (lane == 0) {
      o_[3 + ca] = tra;
      if (hb) o_[3 + cb] = trb;
    }
    // tr(B_c^2) = sum_ab B[a][b] B[b][a] through a transposed read of the tile
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      tile[(lg + 4 * v) * 17 + lr] = A[v];
      tileb[(lg + 4 * v) * 17 + lr] = B[v];
    }
    wave_sync();
#pragma unroll
    for (int v = 0; v < 4; ++v) tsum += A[v] * tile[lr * 17 + lg + 4 * v] + B[v] * tileb[lr * 17 + lg + 4 * v];
    wave_sync();
  }
  tsum = wave_sum_d(tsum);
  if (lane == 0) {
    o_[0] = logdet.log_abs();
    o_[1] = (double)sgn;
    o_[2] = trAL - tsum;
  }
}

// ---------------------------------------------------------------------------
// Per-walker combination: signed log-sum-exp over determinants
// (wavefunction/output/logdet.py:66-79), its gradient and Laplacian, the
// Psiformer Jastrow (wavefunction/jastrow.py:60-122), the Coulomb potential
// (app/molecule/hamiltonian.py:9-22) and E_L = -1/2 (lap + |grad|^2) + V
// (estimator/kinetic/_common.py:61-73).  One wavefront per walker, FP64.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void combine_kernel(FermiDims d, CombineArgs a) {
  const int w = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = d.N, K = d.K, T = 3 * N;
  const int stride = 3 + (a.lapmode ? T : 0);
  const double* det = a.det + (long)w * K * stride;
  const long gw = (long)a.B_off + w;
  const long lw = a.lp_local ? (long)w : gw;  // index of the log|psi| output
  const float* x = a.elec + gw * N * 3;

  double m = -INFINITY;
  for (int k = 0; k < K; ++k) m = fmax(m, det[k * stride]);
  double S = 0.0;
  for (int k = 0; k < K; ++k) S += det[k * stride + 1] * exp(det[k * stride] - m);
  const double sign = S > 0 ? 1.0 : (S < 0 ? -1.0 : 0.0);
  double logpsi = log(fabs(S)) + m;

  // Jastrow: J = sum_{i<j same} -1/4 a^2/(a+r) + sum_{up,dn} -1/2 a^2/(a+r)
  double jv = 0.0, jl = 0.0;
  if (a.jastrow) {
    for (int p = lane; p < N * N; p += 64) {
      const int i = p / N, j = p % N;
      if (j <= i) continue;
      const bool same = (i < d.n_up) == (j < d.n_up);
      const double c = same ? 0.25 : 0.5;
      const double al = same ? a.alpha_par : a.alpha_anti;
      const double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double q = al + r;
      jv += -c * al * al / q;
      const double f1 = c * al * al / (q * q), f2 = -2.0 * c * al * al / (q * q * q);
      jl += 2.0 * (f2 + 2.0 * f1 / r);
    }
    jv = wave_sum_d(jv);
    jl = wave_sum_d(jl);
    logpsi += jv;
  }

  // potential
  double v = 0.0;
  for (int p = lane; p < N * d.M; p += 64) {
    const int i = p / d.M, I = p % d.M;
    const double dx = x[3 * i] - a.atoms[3 * I], dy = x[3 * i + 1] - a.atoms[3 * I + 1],
                 dz = x[3 * i + 2] - a.atoms[3 * I + 2];
    v -= a.charges[I] / sqrt(dx * dx + dy * dy + dz * dz);
  }
  for (int p = lane; p < N * N; p += 64) {
    const int i = p / N, j = p % N;
    if (j <= i) continue;
    const double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
    v += 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
  }
  for (int p = lane; p < d.M * d.M; p += 64) {
    const int I = p / d.M, J = p % d.M;
    if (J <= I) continue;
    const double dx = a.atoms[3 * I] - a.atoms[3 * J], dy = a.atoms[3 * I + 1] - a.atoms[3 * J + 1],
                 dz = a.atoms[3 * I + 2] - a.atoms[3 * J + 2];
    v += (double)a.charges[I] * a.charges[J] / sqrt(dx * dx + dy * dy + dz * dz);
  }
  v = wave_sum_d(v);

  if (!a.lapmode) {
    if (lane == 0) {
      a.logpsi[lw] = (float)logpsi;
      if (a.sign) a.sign[gw] = (float)sign;
    }
    return;
  }
  // weights w_k = s_k exp(l_k - m) / S  (sum_k w_k = 1)
  // grad = sum_k w_k g_k ;  lap = sum_k w_k (lap_k + |g_k|^2) - |grad|^2
  double g2 = 0.0, wl = 0.0;
  for (int c = lane; c < T; c += 64) {
    double g = 0.0;
    for (int k = 0; k < K; ++k) {
      const double wk = det[k * stride + 1] * exp(det[k * stride] - m) / S;
      const double gk = det[k * stride + 3 + c];
      g += wk * gk;
      wl += wk * gk * gk;
    }
    if (a.jastrow) {
      const int i = c / 3, ax = c % 3;
      for (int j = 0; j < N; ++j) {
        if (j == i) continue;
        const bool same = (i < d.n_up) == (j < d.n_up);
        const double cc = same ? 0.25 : 0.5;
        const double al = same ? a.alpha_par : a.alpha_anti;
        const double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double q = al + r;
        const double dax = ax == 0 ? dx : (ax == 1 ? dy : dz);
        g += cc * al * al / (q * q) * dax / r;
      }
    }
    if (a.grad) a.grad[gw * T + c] = (float)g;
    g2 += g * g;
  }
  g2 = wave_sum_d(g2);
  wl = wave_sum_d(wl);
  double lapdet = 0.0;
  for (int k = 0; k < K; ++k) lapdet += det[k * stride + 1] * exp(det[k * stride] - m) / S * det[k * stride + 2];
  // lap log psi = sum_k w_k (lap_k + |g_k|^2) - |grad_det|^2 + lap J;  with the Jastrow
  // the total gradient g already includes grad J, so recompute |grad_det|^2 implicitly:
  // E_kin = -1/2 (lap + |g|^2) where lap = sum_k w_k(lap_k+|g_k|^2) - |g_det|^2 + lapJ and
  // |g|^2 = |g_det + g_J|^2.  Evaluate with the explicit pieces.
  double gdet2 = 0.0, gcross = 0.0;
  if (a.jastrow) {
    for (int c = lane; c < T; c += 64) {
      double gd = 0.0;
      for (int k = 0; k < K; ++k)
        gd += det[k * stride + 1] * exp(det[k * stride] - m) / S * det[k * stride + 3 + c];
      gdet2 += gd * gd;
    }
    gdet2 = wave_sum_d(gdet2);
  } else {
    gdet2 = g2;
  }
  (void)gcross;
  const double lap = lapdet + wl - gdet2 + jl;
  const double ekin = -0.5 * (lap + g2);
  if (lane == 0) {
    a.logpsi[lw] = (float)logpsi;
    if (a.sign) a.sign[gw] = (float)sign;
    a.e_kin[gw] = ekin;
    a.e_pot[gw] = v;
    a.e_loc[gw] = ekin + v;
  }
}

// Orbitals export (MoleculeWavefunction.orbitals): A[w][k][i][o] =
// Phi[unit(w, i)][k*N + o] * env[w][i][k*N + o] (value rows, forward mode).
__global__ void orbital_export_kernel(FermiDims d, const float* __restrict__ Phi, long ldphi,
                                      const float* __restrict__ env, float* __restrict__ out) {
  const int KN = d.K * d.N;
  const long total = (long)d.Bc * d.N * KN;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int ko = (int)(idx % KN);
    const long wi = idx / KN;
    const int i = (int)(wi % d.N), w = (int)(wi / d.N);
    const int k = ko / d.N, o = ko % d.N;
    const float phi = Phi[(long)unit_idx(d, w, i) * d.Rp * ldphi + ko];
    out[(((long)w * d.K + k) * d.N + i) * d.N + o] = phi * env[(wi * KN + ko) * env_stride(d)];
  }
}

}  // namespace

int launch_orbital_export(const FermiDims& d, const float* Phi, long ldphi, const float* env, float* out,
                          hipStream_t s) {
  const long total = (long)d.Bc * d.N * d.K * d.N;
  long g = (total + 255) / 256;
  if (g > 65536L * 8) g = 65536L * 8;
  hipLaunchKernelGGL(orbital_export_kernel, dim3((unsigned)(g < 1 ? 1 : g)), dim3(256), 0, s, d, Phi, ldphi, env,
                     out);
  return 0;
}

int launch_det(const FermiDims& d, const float* Phi, long ldphi, const float* env, double* out, hipStream_t s) {
  if (d.N <= SM_NP) {
    const dim3 grid((unsigned)xcd_grid(d.Bc, (d.K + 3) / 4));
    if (d.Rp > 1)
      hipLaunchKernelGGL(det_small_kernel<true>, grid, dim3(256), 0, s, d, Phi, ldphi, env, out);
    else
      hipLaunchKernelGGL(det_small_kernel<false>, grid, dim3(256), 0, s, d, Phi, ldphi, env, out);
    return 0;
  }
  if (d.Rp > 1 && d.N <= 64) {
    const int np = round_up(d.N, 16);
    const size_t smem = np == 32 ? det_mfma_smem<32>() : np == 48 ? det_mfma_smem<48>() : det_mfma_smem<64>();
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)det_mfma_kernel<32>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      (void)hipFuncSetAttribute((const void*)det_mfma_kernel<48>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      (void)hipFuncSetAttribute((const void*)det_mfma_kernel<64>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      attr = true;
    }
    const dim3 grid((unsigned)xcd_grid(d.Bc, d.K));
    switch (np) {
      case 32: hipLaunchKernelGGL(det_mfma_kernel<32>, grid, dim3(256), smem, s, d, Phi, ldphi, env, out); return 0;
      case 48: hipLaunchKernelGGL(det_mfma_kernel<48>, grid, dim3(256), smem, s, d, Phi, ldphi, env, out); return 0;
      case 64: hipLaunchKernelGGL(det_mfma_kernel<64>, grid, dim3(256), smem, s, d, Phi, ldphi, env, out); return 0;
    }
  }
  if (d.Rp <= 1 && d.N <= 64) {
    const dim3 grid4((unsigned)xcd_grid(d.Bc, (d.K + 3) / 4));
    switch (round_up(d.N, 8)) {
      case 24: hipLaunchKernelGGL(det_fwd_reg_kernel<24>, grid4, dim3(256), 0, s, d, Phi, ldphi, env, out); return 0;
      case 32: hipLaunchKernelGGL(det_fwd_reg_kernel<32>, grid4, dim3(256), 0, s, d, Phi, ldphi, env, out); return 0;
      case 40: hipLaunchKernelGGL(det_fwd_reg_kernel<40>, grid4, dim3(256), 0, s, d, Phi, ldphi, env, out); return 0;
      case 48: hipLaunchKernelGGL(det_fwd_reg_kernel<48>, grid4, dim3(256), 0, s, d, Phi, ldphi, env, out); return 0;
      case 56: hipLaunchKernelGGL(det_fwd_reg_kernel<56>, grid4, dim3(256), 0, s, d, Phi, ldphi, env, out); return 0;
      case 64: hipLaunchKernelGGL(det_fwd_reg_kernel<64>, grid4, dim3(256), 0, s, d, Phi, ldphi, env, out); return 0;
    }
  }
  return -1;  // N > 64
}

int launch_combine(const FermiDims& d, const CombineArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(combine_kernel, dim3(d.Bc), dim3(64), 0, s, d, a);
  return 0;
}
