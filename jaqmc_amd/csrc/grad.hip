// Reverse pass of FermiNet log|psi| with respect to the parameters, seeded per
// walker:  g = sum_w c_w grad_theta log|psi(x_w)|  (the vector-Jacobian product
// jax.value_and_grad forms per walker in LossAndGrad.evaluate_single_walker,
// estimator/loss_grad.py:70-103, already contracted with the walker weights of
// LossAndGrad.finalize_stats, :115-128).
//
// Derivatives restated from the forward definitions:
//   log|psi| = log|sum_k det A_k|:  d/dA_k = w_k A_k^-T,
//            w_k = s_k e^{l_k - m} / sum_j s_j e^{l_j - m}   (output/logdet.py:55-79)
//   A_k = Phi_k * env_k  (elementwise; app/molecule/wavefunction/ferminet.py:93-95)
//   env[k,i,o] = sum_I pi[o,I,k] exp(-a r_iI), a = |sigma| or sigma  (output/envelope.py:124-140)
//   Phi = h_L W_orb per spin channel                               (output/orbital.py:60-78)
//   h_{l+1} = tanh(X_l W_m + g_l W_g + b) (+ h_l, / sqrt 2)        (backbone/ferminet.py:51-63)
//   h2_{l+1} = tanh(h2_l W_2 + b_2) (+ h2_l, / sqrt 2)
//   g_l = spin means of h_l;  X_l = [h_l | means over the first index of h2_l]
// The kernels here are the pieces that are not plain GEMMs: the FP64
// determinant adjoint, the envelope parameter sums, the tanh/residual
// adjoint, the walker sums and the mean adjoints, and the weight-gradient
// contraction X^T dY over walker rows (FP32 MFMA, deterministic slices).

#include "kernels.h"

namespace {

constexpr int GJ_NP = 64;

// dA_k = c_w w_k A_k^-T for one (walker, det) per workgroup, by FP64
// Gauss-Jordan with partial pivoting on [A | I] in LDS.  Writes
//   dPhi[unit(w,i)][k N + o] = dA[i][o] env[i][o]
//   denv[w][i][k N + o]      = dA[i][o] Phi[i][o]
// A walker with c_w = 0 writes zeros without touching its matrices (a zero
// weight stays exact even for a singular walker).
__global__ __launch_bounds__(256) void det_vjp_kernel(FermiDims d, const float* __restrict__ Phi, long ldphi,
                                                      const float* __restrict__ env, const double* __restrict__ det,
                                                      const double* __restrict__ cot, int b0,
                                                      float* __restrict__ dPhi, long lddphi,
                                                      float* __restrict__ denv) {
  __shared__ double M[GJ_NP][2 * GJ_NP + 1];
  __shared__ double red_v[256];
  __shared__ int red_i[256];
  __shared__ double mult[GJ_NP];  // column p of the rows, read before the update
  __shared__ double wk_s;
  const int w = blockIdx.x / d.K, k = blockIdx.x % d.K;
  const int N = d.N, KN = d.K * d.N;
  const int tid = threadIdx.x;
  const double c = cot[b0 + w];
  if (c == 0.0) {
    for (int e = tid; e < N * N; e += 256) {
      const int i = e / N, o = e % N;
      dPhi[(long)unit_idx(d, w, i) * lddphi + k * N + o] = 0.f;
      denv[((long)w * N + i) * KN + k * N + o] = 0.f;
    }
    return;
  }
  if (tid == 0) {
    const double* dw = det + (long)w * d.K * 3;
    double m = -INFINITY;
    for (int j = 0; j < d.K; ++j) m = fmax(m, dw[j * 3]);
    double S = 0.0;
    for (int j = 0; j < d.K; ++j) S += dw[j * 3 + 1] * exp(dw[j * 3] - m);
    wk_s = dw[k * 3 + 1] * exp(dw[k * 3] - m) / S;
  }
  for (int e = tid; e < N * 2 * N; e += 256) {
    const int i = e / (2 * N), j = e % (2 * N);
    double v;
    if (j < N)
      v = (double)Phi[(long)unit_idx(d, w, i) * ldphi + k * N + j] *
          (double)env[(((long)w * N + i) * KN + k * N + j) * env_stride(d)];
    else
      v = (j - N == i) ? 1.0 : 0.0;
    M[i][j] = v;
  }
  __syncthreads();
  for (int p = 0; p < N; ++p) {
    // pivot: largest |M[r][p]| over r >= p
    double best = -1.0;
    int bi = p;
    for (int r = p + tid; r < N; r += 256) {
      const double a = fabs(M[r][p]);
      if (a > best) {
        best = a;
        bi = r;
      }
    }
    red_v[tid] = best;
    red_i[tid] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s && red_v[tid + s] > red_v[tid]) {
        red_v[tid] = red_v[tid + s];
        red_i[tid] = red_i[tid + s];
      }
      __syncthreads();
    }
    const int pr = red_i[0];
    if (pr != p)
      for (int j = tid; j < 2 * N; j += 256) {
        const double t = M[p][j];
        M[p][j] = M[pr][j];
        M[pr][j] = t;
      }
    __syncthreads();
    const double inv = 1.0 / M[p][p];
    __syncthreads();
    for (int j = tid; j < 2 * N; j += 256) M[p][j] *= inv;
    for (int r = tid; r < N; r += 256)
      if (r != p) mult[r] = M[r][p];
    __syncthreads();
    for (int e = tid; e < N * 2 * N; e += 256) {
      const int r = e / (2 * N), j = e % (2 * N);
      if (r != p) M[r][j] -= mult[r] * M[p][j];
    }
    __syncthreads();
  }
  // inverse in columns N..2N-1: Ainv[o][i] = M[o][N + i]; dA[i][o] = c w_k Ainv[o][i]
  const double scale = c * wk_s;
  for (int e = tid; e < N * N; e += 256) {
    const int i = e / N, o = e % N;
    const double da = scale * M[o][N + i];
    const long u = unit_idx(d, w, i);
    const float ph = Phi[u * ldphi + k * N + o];
    const float ev = env[(((long)w * N + i) * KN + k * N + o) * env_stride(d)];
    dPhi[u * lddphi + k * N + o] = (float)(da * ev);
    denv[((long)w * N + i) * KN + k * N + o] = (float)(da * ph);
  }
}

// Envelope parameter gradients, one wavefront per (orbital set, o, I, k),
// its lanes splitting the (walker, electron) terms and summing in FP64 in a
// fixed order:
//   d pi[o,I,k]    += sum_{w, i in set} denv[w][i][k N + o] e_iI
//   d sigma[o,I,k] += sum denv pi e_iI (-r_iI) (sign(sigma) for abs_isotropic)
// accumulated in FP64 into acc (pack layout (N, M, K) per set: pi then sigma,
// at off0 / off1).  The device keeps pi, sigma as [I][k][o].
__global__ void env_grad_kernel(FermiDims d, const float* __restrict__ elec, const float* __restrict__ atoms,
                                const float* __restrict__ pi0, const float* __restrict__ pi1,
                                const float* __restrict__ sg0, const float* __restrict__ sg1, int kind, int nsets,
                                const float* __restrict__ denv, double* __restrict__ acc, long off0, long off1) {
  const int N = d.N, M = d.M, K = d.K, KN = K * N;
  const long per = (long)N * M * K;
  const long tot = per * nsets;
  const int lane = threadIdx.x & 63;
  const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
  for (long t = blockIdx.x * (long)(blockDim.x >> 6) + (threadIdx.x >> 6); t < tot; t += nwaves) {
    const int set = (int)(t / per);
    const long q = t % per;  // (o, I, k) in pack order
    const int o = (int)(q / ((long)M * K)), I = (int)((q / K) % M), k = (int)(q % K);
    const long qd = ((long)I * K + k) * N + o;  // device layout
    const float pi = (set ? pi1 : pi0)[qd];
    const float sg = (set ? sg1 : sg0)[qd];
    const float a = kind == 1 ? fabsf(sg) : sg;
    const float sgn = kind == 1 ? (sg > 0.f ? 1.f : (sg < 0.f ? -1.f : 0.f)) : 1.f;
    int i0 = 0, i1 = N;
    if (nsets == 2) {
      i0 = set == 0 ? 0 : d.n_up;
      i1 = set == 0 ? d.n_up : N;
    }
    double gp = 0.0, gs = 0.0;
    const int ni = i1 - i0;
    for (long wi = lane; wi < (long)d.Bc * ni; wi += 64) {
      const int w = (int)(wi / ni), i = i0 + (int)(wi % ni);
      const float* x = elec + ((long)w * N + i) * 3;
      const float dx = x[0] - atoms[3 * I], dy = x[1] - atoms[3 * I + 1], dz = x[2] - atoms[3 * I + 2];
      const float r = sqrtf(dx * dx + dy * dy + dz * dz);
      const float e = expf(-a * r);
      const double de = denv[((long)w * N + i) * KN + k * N + o];
      gp += de * e;
      gs += de * (double)(pi * e * (-r) * sgn);
    }
    gp = wave_sum_d(gp);
    gs = wave_sum_d(gs);
    if (lane == 0) {
      const long off = set ? off1 : off0;
      acc[off + q] += gp;
      acc[off + per + q] += gs;
    }
  }
}

// tanh and residual adjoint of h_new = tanh(z) (+ h_old) / sqrt 2:
//   t  = res ? sqrt2 h_new - h_old : h_new,   dz = dh_new s (1 - t^2)
//   dh_old = res ? dh_new / sqrt 2 : 0   (initialises the next adjoint)
// s = 1/sqrt 2 with the residual, 1 without.
__global__ void tanh_bwd_kernel(long rows, int n, const float* __restrict__ hnew, long ldn,
                                const float* __restrict__ hold, long ldo, const float* __restrict__ dhnew, long lddn,
                                int res, float* __restrict__ dz, long lddz, float* __restrict__ dhold, long lddo) {
  const long tot = rows * n;
  const float r2 = 0.70710678118654752f;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long r = e / n;
    const int c = (int)(e % n);
    const float hn = hnew[r * ldn + c];
    const float g = dhnew[r * lddn + c];
    float t = hn;
    if (res) t = 1.41421356237309515f * hn - hold[r * ldo + c];
    dz[r * lddz + c] = (res ? g * r2 : g) * (1.f - t * t);
    if (dhold) dhold[r * lddo + c] = res ? g * r2 : 0.f;
  }
}

// dG[w][n] = sum over the walker's units of dz[unit][n]
__global__ void walker_sum_kernel(FermiDims d, const float* __restrict__ dz, long lddz, int n,
                                  float* __restrict__ dG, long lddg) {
  const long tot = (long)d.Bc * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const int w = (int)(e / n), c = (int)(e % n);
    float s = 0.f;
    for (int i = 0; i < d.N; ++i) s += dz[(long)unit_idx(d, w, i) * lddz + c];
    dG[(long)w * lddg + c] = s;
  }
}

// dh[unit(w,i)][f] += dX[unit(w,i)][f] + dg[w][ch(i) f1 + f] / n_ch(i)   (f < f1)
__global__ void one_adj_kernel(FermiDims d, int f1, const float* __restrict__ dX, long lddx,
                               const float* __restrict__ dg, long lddg, float* __restrict__ dh, long lddh) {
  const long tot = (long)d.Bc * d.N * f1;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const int f = (int)(e % f1);
    const long wi = e / f1;
    const int w = (int)(wi / d.N), i = (int)(wi % d.N);
    const int ch = (d.nch == 2 && i >= d.n_up) ? 1 : 0;
    const int nc = d.nch == 2 ? (ch ? d.n_dn : d.n_up) : d.N;
    const long u = unit_idx(d, w, i);
    dh[u * lddh + f] += dX[u * lddx + f] + dg[(long)w * lddg + ch * f1 + f] / (float)nc;
  }
}

// dh2[w][j][i][f] += dX[unit(w,i)][col0 + ch(j) f2 + f] / n_ch(j)  (pair means
// over the first index j within its spin channel)
__global__ void pair_adj_kernel(FermiDims d, int f2, const float* __restrict__ dX, long lddx, int col0,
                                float* __restrict__ dh2) {
  const long tot = (long)d.Bc * d.N * d.N * f2;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const int f = (int)(e % f2);
    const long p = e / f2;
    const int i = (int)(p % d.N), j = (int)((p / d.N) % d.N), w = (int)(p / ((long)d.N * d.N));
    const int ch = (d.nch == 2 && j >= d.n_up) ? 1 : 0;
    const int nc = d.nch == 2 ? (ch ? d.n_dn : d.n_up) : d.N;
    dh2[e] += dX[(long)unit_idx(d, w, i) * lddx + col0 + ch * f2 + f] / (float)nc;
  }
}

// dh2_old[p][f] = sum_n dz2[p][n] W2[f][n] (+ dh2_new[p][f] / sqrt 2 with the residual)
__global__ void pair_dx_kernel(long npairs, int fin, int fout, const float* __restrict__ dz2,
                               const float* __restrict__ W2, int res, const float* __restrict__ dh2new,
                               float* __restrict__ dh2old) {
  const long tot = npairs * fin;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long p = e / fin;
    const int f = (int)(e % fin);
    float s = 0.f;
    for (int n = 0; n < fout; ++n) s += dz2[p * fout + n] * W2[f * fout + n];
    if (res) s += dh2new[p * fout + f] * 0.70710678118654752f;
    dh2old[e] = s;
  }
}

// Weight-gradient contraction over rows on FP32 MFMA (v_mfma_f32_16x16x4f32):
//   P[s][k][n] = sum_{r in slice s} X[r][k] dY[r][n]   (k < Kx, n < Ny)
// X null: a column of ones (k = 0 only: column sums, the bias gradients).
// The A operand of a 16x16x4 step is X^T (row m = feature k, inner = row r),
// the B operand dY: lane l reads X[r0 + l/16][k0 + l%16] and dY[r0 + l/16][n0 +
// l%16], 64-byte row segments per lane group, no LDS staging.  A workgroup
// (4 waves) owns a 64 x 64 output tile of one slice; the waves take every
// fourth group of 4 rows and their partial tiles are summed through LDS.
constexpr int WG_T = 4;  // 16x16 blocks per wave tile side

__global__ __launch_bounds__(256) void wgrad_kernel(long rows, long rows_per_slice, const float* __restrict__ X,
                                                    long ldx, int Kx, const float* __restrict__ dY, long ldy, int Ny,
                                                    float* __restrict__ P) {
  __shared__ float red[64 * 65];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int k0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const long r_beg = (long)blockIdx.z * rows_per_slice;
  const long r_end = r_beg + rows_per_slice < rows ? r_beg + rows_per_slice : rows;
  f32x4 acc[WG_T][WG_T];
#pragma unroll
  for (int a = 0; a < WG_T; ++a)
#pragma unroll
    for (int b = 0; b < WG_T; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long r0 = r_beg + 4 * wave; r0 < r_end; r0 += 16) {
    const long r = r0 + lg;
    const bool rok = r < r_end;
    float av[WG_T], bv[WG_T];
#pragma unroll
    for (int a = 0; a < WG_T; ++a) {
      const int k = k0 + a * 16 + lr;
      av[a] = (rok && k < Kx) ? (X ? X[r * ldx + k] : (k == 0 ? 1.f : 0.f)) : 0.f;
      const int n = n0 + a * 16 + lr;
      bv[a] = (rok && n < Ny) ? dY[r * ldy + n] : 0.f;
    }
#pragma unroll
    for (int a = 0; a < WG_T; ++a)
#pragma unroll
      for (int b = 0; b < WG_T; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
  // acc[a][b][v]: output row k0 + a*16 + 4*lg + v, column n0 + b*16 + lr
  for (int src = 0; src < 4; ++src) {
    if (wave == src) {
#pragma unroll
      for (int a = 0; a < WG_T; ++a)
#pragma unroll
        for (int b = 0; b < WG_T; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            float* t = &red[(a * 16 + 4 * lg + v) * 65 + b * 16 + lr];
            *t = (src == 0 ? 0.f : *t) + acc[a][b][v];
          }
    }
    __syncthreads();
  }
  float* Ps = P + (long)blockIdx.z * Kx * Ny;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int kk = e / 64, nn = e % 64;
    if (k0 + kk < Kx && n0 + nn < Ny) Ps[(long)(k0 + kk) * Ny + n0 + nn] = red[kk * 65 + nn];
  }
}

// acc[off + (row0 + k) * ld + n] += sum_s P[s][k][n]   (FP64, fixed slice order)
__global__ void slice_sum_kernel(const float* __restrict__ P, int ns, int Kx, int Ny, double* __restrict__ acc,
                                 long off, int row0, int ld) {
  const long tot = (long)Kx * Ny;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < ns; ++z) s += P[(long)z * tot + e];
    const int k = (int)(e / Ny), n = (int)(e % Ny);
    acc[off + (long)(row0 + k) * ld + n] += s;
  }
}

__global__ void fill_f64_kernel(double* p, long n, double v) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = v;
}
__global__ void fill_f32_kernel(float* p, long n, float v) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = v;
}
__global__ void f64_to_f32_kernel(const double* a, long n, float* out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    out[e] = (float)a[e];
}

inline int grid_of(long total) {
  long g = (total + 255) / 256;
  if (g > 65536L * 4) g = 65536L * 4;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

int launch_det_vjp(const FermiDims& d, const float* Phi, long ldphi, const float* env, const double* det,
                   const double* cot, int b0, float* dPhi, long lddphi, float* denv, hipStream_t s) {
  if (d.N > GJ_NP) return -1;
  hipLaunchKernelGGL(det_vjp_kernel, dim3(d.Bc * d.K), dim3(256), 0, s, d, Phi, ldphi, env, det, cot, b0, dPhi,
                     lddphi, denv);
  return 0;
}

int launch_env_grad(const FermiDims& d, const float* elec, const float* atoms, const float* pi0, const float* pi1,
                    const float* sg0, const float* sg1, int kind, int nsets, const float* denv, double* acc,
                    long off0, long off1, hipStream_t s) {
  const long tot = (long)d.N * d.M * d.K * nsets;  // one wavefront each
  hipLaunchKernelGGL(env_grad_kernel, dim3(grid_of(tot * 64)), dim3(256), 0, s, d, elec, atoms, pi0, pi1, sg0, sg1,
                     kind, nsets, denv, acc, off0, off1);
  return 0;
}

int launch_tanh_bwd(long rows, int n, const float* hnew, long ldn, const float* hold, long ldo, const float* dhnew,
                    long lddn, int res, float* dz, long lddz, float* dhold, long lddo, hipStream_t s) {
  hipLaunchKernelGGL(tanh_bwd_kernel, dim3(grid_of(rows * n)), dim3(256), 0, s, rows, n, hnew, ldn, hold, ldo, dhnew,
                     lddn, res, dz, lddz, dhold, lddo);
  return 0;
}

int launch_walker_sum(const FermiDims& d, const float* dz, long lddz, int n, float* dG, long lddg, hipStream_t s) {
  hipLaunchKernelGGL(walker_sum_kernel, dim3(grid_of((long)d.Bc * n)), dim3(256), 0, s, d, dz, lddz, n, dG, lddg);
  return 0;
}

int launch_one_adj(const FermiDims& d, int f1, const float* dX, long lddx, const float* dg, long lddg, float* dh,
                   long lddh, hipStream_t s) {
  hipLaunchKernelGGL(one_adj_kernel, dim3(grid_of((long)d.Bc * d.N * f1)), dim3(256), 0, s, d, f1, dX, lddx, dg,
                     lddg, dh, lddh);
  return 0;
}

int launch_pair_adj(const FermiDims& d, int f2, const float* dX, long lddx, int col0, float* dh2, hipStream_t s) {
  hipLaunchKernelGGL(pair_adj_kernel, dim3(grid_of((long)d.Bc * d.N * d.N * f2)), dim3(256), 0, s, d, f2, dX, lddx,
                     col0, dh2);
  return 0;
}

int launch_pair_dx(long npairs, int fin, int fout, const float* dz2, const float* W2, int res, const float* dh2new,
                   float* dh2old, hipStream_t s) {
  hipLaunchKernelGGL(pair_dx_kernel, dim3(grid_of(npairs * fin)), dim3(256), 0, s, npairs, fin, fout, dz2, W2, res,
                     dh2new, dh2old);
  return 0;
}

// P needs nslices * Kx * Ny floats; returns the slice count used.
int wgrad_slices(long rows, int Kx, int Ny, long pcap) {
  const long tiles = (long)((Kx + 63) / 64) * ((Ny + 63) / 64);
  long ns = 1024 / (tiles > 0 ? tiles : 1);
  ns = ns < 1 ? 1 : (ns > 256 ? 256 : ns);
  const long by_rows = (rows + 63) / 64;
  if (ns > by_rows) ns = by_rows > 0 ? by_rows : 1;
  while (ns > 1 && ns * Kx * Ny > pcap) ns >>= 1;
  return (int)ns;
}

int launch_wgrad(long rows, const float* X, long ldx, int Kx, const float* dY, long ldy, int Ny, float* P, long pcap,
                 double* acc, long off, int row0, int ld, hipStream_t s) {
  if (rows <= 0 || Kx <= 0 || Ny <= 0) return 0;
  const int ns = wgrad_slices(rows, Kx, Ny, pcap);
  if ((long)ns * Kx * Ny > pcap) return -1;
  long rps = (rows + ns - 1) / ns;
  rps = (rps + 3) / 4 * 4;
  const dim3 grid((Kx + 63) / 64, (Ny + 63) / 64, ns);
  hipLaunchKernelGGL(wgrad_kernel, grid, dim3(256), 0, s, rows, rps, X, ldx, Kx, dY, ldy, Ny, P);
  hipLaunchKernelGGL(slice_sum_kernel, dim3(grid_of((long)Kx * Ny)), dim3(256), 0, s, P, ns, Kx, Ny, acc, off, row0,
                     ld);
  return 0;
}

int launch_fill_f64(double* p, long n, double v, hipStream_t s) {
  hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_of(n)), dim3(256), 0, s, p, n, v);
  return 0;
}
int launch_fill_f32(float* p, long n, float v, hipStream_t s) {
  hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_of(n)), dim3(256), 0, s, p, n, v);
  return 0;
}
int launch_f64_to_f32(const double* a, long n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_of(n)), dim3(256), 0, s, a, n, out);
  return 0;
}
