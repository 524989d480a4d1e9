// Metropolis-Hastings kernels and FP64 walker reductions.
//
// Proposal (sampler/mcmc.py:53-54): x' = x + width * xi, xi ~ N(0, 1) per
// coordinate.  Accept (sampler/mcmc.py:127-137): ratio = logp(x') - logp(x),
// accept iff ratio > log(u) (strict), u ~ U(0, 1).
//
// Randomness, when not supplied by the caller, is Philox-4x32-10 (Salmon et al.,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) with key = seed and
// counter = (block index, walker, step low word, step high word ^ stream tag);
// normals by the Box-Muller transform.  Streams are keyed by (seed, step,
// walker, element), never by launch geometry, so a walker's trajectory does
// not depend on the batch it is evaluated in.

#include "kernels.h"

namespace {

struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ inline uint32_t mulhilo(uint32_t a, uint32_t b, uint32_t& hi) {
  const uint64_t p = (uint64_t)a * b;
  hi = (uint32_t)(p >> 32);
  return (uint32_t)p;
}

__device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, hi1;
    const uint32_t lo0 = mulhilo(0xD2511F53u, c.x, hi0);
    const uint32_t lo1 = mulhilo(0xCD9E8D57u, c.z, hi1);
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// (0, 1], never 0 so log(u) is finite
__device__ inline float u01(uint32_t v) { return ((v >> 8) + 1u) * (1.0f / 16777216.0f); }

constexpr uint32_t TAG_NORMAL = 0u;
constexpr uint32_t TAG_UNIFORM = 0x80000000u;

__global__ void propose_kernel(const float* __restrict__ elec, float* __restrict__ prop, int B, int N3, float width,
                               const double* __restrict__ width_dev, const float* __restrict__ normals,
                               uint64_t seed, uint64_t step, int w0) {
  // one thread per group of 4 coordinates of one walker
  if (width_dev) width = (float)width_dev[0];
  const int groups = (N3 + 3) / 4;
  const long total = (long)B * groups;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int gi = idx % groups;
    const int w = idx / groups;
    float z[4];
    if (normals) {
      for (int e = 0; e < 4; ++e) {
        const int c = 4 * gi + e;
        z[e] = c < N3 ? normals[(long)w * N3 + c] : 0.f;
      }
    } else {
      const u32x4 r = philox4x32_10(
          u32x4{(uint32_t)gi, (uint32_t)(w0 + w), (uint32_t)step, (uint32_t)(step >> 32) ^ TAG_NORMAL},
          (uint32_t)seed, (uint32_t)(seed >> 32));
      const float u1 = u01(r.x), u2 = u01(r.y), u3 = u01(r.z), u4 = u01(r.w);
      const float m1 = sqrtf(-2.f * logf(u1)), m2 = sqrtf(-2.f * logf(u3));
      float s1, c1, s2, c2;
      sincospif(2.f * u2, &s1, &c1);
      sincospif(2.f * u4, &s2, &c2);
      z[0] = m1 * c1;
      z[1] = m1 * s1;
      z[2] = m2 * c2;
      z[3] = m2 * s2;
    }
    for (int e = 0; e < 4; ++e) {
      const int c = 4 * gi + e;
      if (c < N3) prop[(long)w * N3 + c] = elec[(long)w * N3 + c] + width * z[e];
    }
  }
}

__global__ void accept_kernel(float* __restrict__ elec, const float* __restrict__ prop, float* __restrict__ logp,
                              const float* __restrict__ logpsi_prop, int B, int N3, const float* __restrict__ uniforms,
                              uint64_t seed, uint64_t step, int w0, int32_t* __restrict__ n_accept) {
  const int w = blockIdx.x;
  __shared__ int acc_s;
  if (threadIdx.x == 0) {
    float u;
    if (uniforms) {
      u = uniforms[w];
    } else {
      const u32x4 r = philox4x32_10(
          u32x4{0u, (uint32_t)(w0 + w), (uint32_t)step, (uint32_t)(step >> 32) ^ TAG_UNIFORM},
          (uint32_t)seed, (uint32_t)(seed >> 32));
      u = u01(r.x);
    }
    const float lp2 = 2.f * logpsi_prop[w];
    const float ratio = lp2 - logp[w];
    const int acc = ratio > logf(u) ? 1 : 0;
    if (acc) logp[w] = lp2;
    n_accept[w] += acc;
    acc_s = acc;
  }
  __syncthreads();
  if (acc_s)
    for (int c = threadIdx.x; c < N3; c += blockDim.x) elec[(long)w * N3 + c] = prop[(long)w * N3 + c];
}

// nanmean / variance over finite entries in FP64 (estimator/base.py:27-53):
// one workgroup, Kahan-free two-level reduction (values are O(1e3) at most and
// B <= 1e7, far inside FP64's range for a plain sum).
__global__ __launch_bounds__(1024) void mean_var_kernel(const double* __restrict__ x, int B, double* __restrict__ out) {
  __shared__ double s1[16], s2[16], s3[16];
  double a = 0.0, b = 0.0, n = 0.0;
  for (int i = threadIdx.x; i < B; i += blockDim.x) {
    const double v = x[i];
    if (isfinite(v)) {
      a += v;
      b += v * v;
      n += 1.0;
    }
  }
  a = wave_sum_d(a);
  b = wave_sum_d(b);
  n = wave_sum_d(n);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s1[wv] = a;
    s2[wv] = b;
    s3[wv] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double A = 0, Bs = 0, Nn = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      A += s1[i];
      Bs += s2[i];
      Nn += s3[i];
    }
    const double m = A / Nn;
    out[0] = m;
    out[1] = Bs / Nn - m * m;
    out[2] = Nn;
  }
}

__global__ __launch_bounds__(1024) void sum_i32_kernel(const int32_t* __restrict__ x, int B, double* __restrict__ out) {
  __shared__ double s1[16];
  double a = 0.0;
  for (int i = threadIdx.x; i < B; i += blockDim.x) a += (double)x[i];
  a = wave_sum_d(a);
  if ((threadIdx.x & 63) == 0) s1[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double A = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) A += s1[i];
    out[0] = A;
  }
}

// MCMCSampler.step's bookkeeping after `steps` moves (sampler/mcmc.py:179-195)
// on the device state {stddev, counter, accepted total, pmoves[af]}:
// pmove = (sum of n_accept - previous total) / (steps * B); counter += 1;
// pmoves[counter % af] = pmove; when counter % af == 0 the width is multiplied
// by 1.1 if mean(pmoves) > hi, divided by 1.1 if < lo.  One workgroup.
__global__ __launch_bounds__(1024) void mcmc_adapt_kernel(double* __restrict__ st, const int32_t* __restrict__ n_accept,
                                                          int B, int steps, int af, double lo, double hi,
                                                          double* __restrict__ pmove_out) {
  __shared__ double s1[16];
  double a = 0.0;
  for (int i = threadIdx.x; i < B; i += blockDim.x) a += (double)n_accept[i];
  a = wave_sum_d(a);
  if ((threadIdx.x & 63) == 0) s1[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) tot += s1[i];
    const double pmove = (tot - st[2]) / ((double)steps * (double)B);
    st[2] = tot;
    const double counter = st[1] + 1.0;
    st[1] = counter;
    const long t = (long)counter % af;
    st[3 + t] = pmove;
    if (t == 0) {
      double m = 0.0;
      for (int i = 0; i < af; ++i) m += st[3 + i];
      m /= af;
      if (m > hi)
        st[0] = st[0] * 1.1;
      else if (m < lo)
        st[0] = st[0] / 1.1;
    }
    if (pmove_out) pmove_out[0] = pmove;
  }
}

__global__ void scale_kernel(float* x, int n, float a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

}  // namespace

int launch_propose(const float* elec, float* prop, int B, int N, float width, const double* width_dev,
                   const float* normals, uint64_t seed, uint64_t step, int w0, hipStream_t s) {
  const long total = (long)B * ((3 * N + 3) / 4);
  const int grid = (int)((total + 255) / 256);
  hipLaunchKernelGGL(propose_kernel, grid, 256, 0, s, elec, prop, B, 3 * N, width, width_dev, normals, seed, step,
                     w0);
  return 0;
}

int launch_accept(float* elec, const float* prop, float* logp, const float* logpsi_prop, int B, int N,
                  const float* uniforms, uint64_t seed, uint64_t step, int w0, int32_t* n_accept, hipStream_t s) {
  hipLaunchKernelGGL(accept_kernel, B, 64, 0, s, elec, prop, logp, logpsi_prop, B, 3 * N, uniforms, seed, step, w0,
                     n_accept);
  return 0;
}

int launch_mcmc_adapt(double* state, const int32_t* n_accept, int B, int steps, int af, double lo, double hi,
                      double* pmove_out, hipStream_t s) {
  hipLaunchKernelGGL(mcmc_adapt_kernel, 1, 1024, 0, s, state, n_accept, B, steps, af, lo, hi, pmove_out);
  return 0;
}

int launch_reduce_mean_var(const double* x, int B, double* out, hipStream_t s) {
  hipLaunchKernelGGL(mean_var_kernel, 1, 1024, 0, s, x, B, out);
  return 0;
}

int launch_reduce_accepts(const int32_t* x, int B, double* out, hipStream_t s) {
  hipLaunchKernelGGL(sum_i32_kernel, 1, 1024, 0, s, x, B, out);
  return 0;
}

int launch_scale_f32(float* x, int n, float a, hipStream_t s) {
  hipLaunchKernelGGL(scale_kernel, (n + 255) / 256, 256, 0, s, x, n, a);
  return 0;
}
