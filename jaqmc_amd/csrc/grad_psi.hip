// Reverse pass of the Psiformer value path with respect to the parameters:
// the pieces that are not plain GEMMs (the dense layers' adjoints and weight
// gradients run on the GEMM family and wgrad_kernel, grad.hip).  Derivatives
// restated from the forward definitions (wavefunction/backbone/psiformer.py:
// 29-98, wavefunction/jastrow.py:60-122):
//   LayerNorm  y = s xh + b,  xh = (x - mu) / sigma,  sigma = sqrt(var + eps):
//     g = dy s,  dx = (g - mean(g) - xh mean(g xh)) / sigma,
//     ds = sum_rows dy xh,  db = sum_rows dy
//   attention per (walker, head), q carrying 1/sqrt(dh):  a = softmax_j(q.k),
//   o = a v:  dv = a^T do,  da = do v^T,  ds = a (da - rowsum(a da)),
//     dq = ds k,  dk = ds^T q
//   Jastrow  -c alpha^2 / (alpha + r):  d/d alpha = -c alpha (alpha + 2 r) / (alpha + r)^2

#include "kernels.h"

namespace {

constexpr int LN_MAXV = 8;  // features per lane (D <= 512)

// One wavefront per row.  out = base + dx (base may alias out; null: out =
// dx); P = dy xh (the rows of the scale gradient).
__global__ __launch_bounds__(256) void ln_bwd_kernel(long rows, int D, const float* __restrict__ x, long ldx,
                                                     const float* __restrict__ s, const float* __restrict__ dy,
                                                     long lddy, float eps, const float* base, long ldb, float* out,
                                                     long ldo, float* __restrict__ P, long ldp) {
  const long r = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  float xv[LN_MAXV], gv[LN_MAXV], dv[LN_MAXV];
  float sum = 0.f;
#pragma unroll
  for (int q = 0; q < LN_MAXV; ++q) {
    const int c = lane + 64 * q;
    xv[q] = c < D ? x[r * ldx + c] : 0.f;
    dv[q] = c < D ? dy[r * lddy + c] : 0.f;
    sum += xv[q];
  }
  const float mu = wave_sum(sum) / (float)D;
  float sq = 0.f;
#pragma unroll
  for (int q = 0; q < LN_MAXV; ++q) {
    const float t = lane + 64 * q < D ? xv[q] - mu : 0.f;
    sq += t * t;
  }
  const float rs = 1.f / sqrtf(wave_sum(sq) / (float)D + eps);
  float sg = 0.f, sgx = 0.f;
#pragma unroll
  for (int q = 0; q < LN_MAXV; ++q) {
    const int c = lane + 64 * q;
    const float xh = c < D ? (xv[q] - mu) * rs : 0.f;
    const float g = c < D ? dv[q] * s[c] : 0.f;
    xv[q] = xh;
    gv[q] = g;
    sg += g;
    sgx += g * xh;
  }
  const float mg = wave_sum(sg) / (float)D, mgx = wave_sum(sgx) / (float)D;
#pragma unroll
  for (int q = 0; q < LN_MAXV; ++q) {
    const int c = lane + 64 * q;
    if (c < D) {
      out[r * ldo + c] = (base ? base[r * ldb + c] : 0.f) + (gv[q] - mg - xv[q] * mgx) * rs;
      P[r * ldp + c] = dv[q] * xv[q];
    }
  }
}

// One workgroup per (walker, head).  Q, K, V come head-split
// ([w][part][h][i][dh], one row per electron), dO row-major with the head's
// features at columns h*dh; dq, dk, dv go to row-major [unit][part*D + h*dh + d].
__global__ __launch_bounds__(256) void attn_bwd_kernel(const float* __restrict__ QKV, const float* __restrict__ dO,
                                                       long lddo, float* __restrict__ dQKV, long ldq, int N, int H,
                                                       int dh) {
  extern __shared__ float sm[];
  const int SR = dh + 1, SN = N + 1;
  float* Qs = sm;
  float* Ks = Qs + N * SR;
  float* Vs = Ks + N * SR;
  float* Gs = Vs + N * SR;  // dO
  float* As = Gs + N * SR;  // scores, then a
  float* Ds = As + N * SN;  // da, then ds
  const int w = blockIdx.x / H, h = blockIdx.x % H, tid = threadIdx.x;
  const long blk = (long)N * dh;
  const float* q = QKV + ((long)w * 3 * H + h) * blk;
  for (int e = tid; e < N * dh; e += 256) {
    const int i = e / dh, d = e % dh;
    Qs[i * SR + d] = q[e];
    Ks[i * SR + d] = q[(long)H * blk + e];
    Vs[i * SR + d] = q[2L * H * blk + e];
    Gs[i * SR + d] = dO[((long)w * N + i) * lddo + h * dh + d];
  }
  __syncthreads();
  for (int e = tid; e < N * N; e += 256) {
    const int i = e / N, j = e % N;
    float sa = 0.f, da = 0.f;
    for (int d = 0; d < dh; ++d) {
      sa += Qs[i * SR + d] * Ks[j * SR + d];
      da += Gs[i * SR + d] * Vs[j * SR + d];
    }
    As[i * SN + j] = sa;
    Ds[i * SN + j] = da;
  }
  __syncthreads();
  for (int i = tid; i < N; i += 256) {
    float m = -INFINITY;
    for (int j = 0; j < N; ++j) m = fmaxf(m, As[i * SN + j]);
    float z = 0.f;
    for (int j = 0; j < N; ++j) {
      const float e = expf(As[i * SN + j] - m);
      As[i * SN + j] = e;
      z += e;
    }
    const float iz = 1.f / z;
    float rsum = 0.f;
    for (int j = 0; j < N; ++j) {
      const float a = As[i * SN + j] * iz;
      As[i * SN + j] = a;
      rsum += a * Ds[i * SN + j];
    }
    for (int j = 0; j < N; ++j) Ds[i * SN + j] = As[i * SN + j] * (Ds[i * SN + j] - rsum);
  }
  __syncthreads();
  const int D = H * dh;
  float* out = dQKV + (long)w * N * ldq + h * dh;
  for (int e = tid; e < N * dh; e += 256) {
    const int i = e / dh, d = e % dh;
    float gq = 0.f, gk = 0.f, gv = 0.f;
    for (int j = 0; j < N; ++j) {
      gq += Ds[i * SN + j] * Ks[j * SR + d];
      gk += Ds[j * SN + i] * Qs[j * SR + d];
      gv += As[j * SN + i] * Gs[j * SR + d];
    }
    out[(long)i * ldq + d] = gq;
    out[(long)i * ldq + D + d] = gk;
    out[(long)i * ldq + 2 * D + d] = gv;
  }
}

// out[w][0..1] = c_w dJ_w / d(alpha_par, alpha_anti), one thread per walker
// of the chunk (walker b0 + w of the weights)
__global__ void jastrow_grad_kernel(const float* __restrict__ elec, int Bc, int N, int n_up, float ap, float aa,
                                    const double* __restrict__ cot, int b0, float* __restrict__ out) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= Bc) return;
  const float* x = elec + (long)w * N * 3;
  double gp = 0.0, ga = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = i + 1; j < N; ++j) {
      const double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if ((i < n_up) == (j < n_up)) {
        const double q = ap + r;
        gp += -0.25 * ap * (ap + 2.0 * r) / (q * q);
      } else {
        const double q = aa + r;
        ga += -0.5 * aa * (aa + 2.0 * r) / (q * q);
      }
    }
  const double c = cot[b0 + w];
  out[2 * w] = (float)(c * gp);
  out[2 * w + 1] = (float)(c * ga);
}

// dst[unit][n] = src[unit][n] for the electrons i0 <= i < i1 of every walker
// (walker-major units), 0 for the others
__global__ void mask_rows_kernel(long units, int N, int i0, int i1, int ncols, const float* __restrict__ src, long lds,
                                 float* __restrict__ dst, long ldd) {
  const long tot = units * ncols;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long u = e / ncols;
    const int n = (int)(e % ncols), i = (int)(u % N);
    dst[u * ldd + n] = (i >= i0 && i < i1) ? src[u * lds + n] : 0.f;
  }
}

// out[r][c] = a[r][c] + b[r][c]
__global__ void add_rows_kernel(long rows, int n, const float* a, long lda, const float* b, long ldb, float* out,
                                long ldo) {
  const long tot = rows * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long r = e / n;
    const int c = (int)(e % n);
    out[r * ldo + c] = a[r * lda + c] + b[r * ldb + c];
  }
}

__global__ void scale_f64_kernel(double* p, long n, double v) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] *= v;
}

inline int grid_for(long total) {
  long g = (total + 255) / 256;
  if (g > 65536L * 4) g = 65536L * 4;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

int launch_ln_bwd(long rows, int D, const float* x, long ldx, const float* s, const float* dy, long lddy, float eps,
                  const float* base, long ldb, float* out, long ldo, float* P, long ldp, hipStream_t st) {
  if (D > 64 * LN_MAXV || rows <= 0) return rows <= 0 ? 0 : -1;
  hipLaunchKernelGGL(ln_bwd_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, st, rows, D, x, ldx, s, dy, lddy,
                     eps, base, ldb, out, ldo, P, ldp);
  return 0;
}

size_t attn_bwd_smem(int N, int dh) { return (size_t)(4 * N * (dh + 1) + 2 * N * (N + 1)) * 4; }

int launch_attn_bwd(const float* QKV, const float* dO, long lddo, float* dQKV, long ldq, int Bc, int N, int H, int dh,
                    hipStream_t st) {
  const size_t smem = attn_bwd_smem(N, dh);
  if (N > 64 || dh > 64 || Bc <= 0) return Bc <= 0 ? 0 : -1;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)attn_bwd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)attn_bwd_smem(64, 64));
    attr = true;
  }
  hipLaunchKernelGGL(attn_bwd_kernel, dim3((unsigned)(Bc * H)), dim3(256), smem, st, QKV, dO, lddo, dQKV, ldq, N, H,
                     dh);
  return 0;
}

int launch_jastrow_grad(const float* elec, int Bc, int N, int n_up, float ap, float aa, const double* cot, int b0,
                        float* out, hipStream_t st) {
  if (Bc <= 0) return 0;
  hipLaunchKernelGGL(jastrow_grad_kernel, dim3((unsigned)((Bc + 63) / 64)), dim3(64), 0, st, elec, Bc, N, n_up, ap, aa,
                     cot, b0, out);
  return 0;
}

int launch_mask_rows(long units, int N, int i0, int i1, int ncols, const float* src, long lds, float* dst, long ldd,
                     hipStream_t st) {
  hipLaunchKernelGGL(mask_rows_kernel, dim3(grid_for(units * ncols)), dim3(256), 0, st, units, N, i0, i1, ncols, src,
                     lds, dst, ldd);
  return 0;
}

int launch_add_rows(long rows, int n, const float* a, long lda, const float* b, long ldb, float* out, long ldo,
                    hipStream_t st) {
  hipLaunchKernelGGL(add_rows_kernel, dim3(grid_for(rows * n)), dim3(256), 0, st, rows, n, a, lda, b, ldb, out, ldo);
  return 0;
}

int launch_scale_f64(double* p, long n, double v, hipStream_t st) {
  hipLaunchKernelGGL(scale_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, n, v);
  return 0;
}
