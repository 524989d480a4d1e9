// Outlier clipping of per-walker losses for the VMC gradient
// (estimator/loss_grad.py:91-128, utils/clip.py:11-60), on the device in FP64:
//   "mad":  median m = nanmedian(x), d = nanmedian(|x - m|), clip to m -+ s d
//   "iqr":  q1, q3 = nanquantile(x, .25 / .75), clip to [q1 - s iqr, q3 + s iqr]
//   "none": no clipping
// then the gradient weights  w_i = 2 (c_i - mean(c)) / n  over the n walkers
// whose clipped loss c_i is not NaN (0 for the others), so that
//   sum_i w_i grad log|psi_i| = 2 (mean(grad.c) - mean(c) mean(grad)),
// LossAndGrad.finalize_stats's gradient.
//
// Quantiles follow jnp.nanquantile's default ("linear"): position q (n - 1)
// between the order statistics floor and ceil of it, NaN entries ignored.  An order
// statistic is found by an MSB-first radix select over the FP64 values mapped
// to order-preserving 64-bit keys, 8 bits per pass: each pass histograms the
// digit of the keys that match the prefix found so far, one partial histogram
// per workgroup into the scratch buffer, and one workgroup sums the partials
// and picks the digit holding the target rank.  Nothing in the selection
// depends on the workgroup count or order, so results are deterministic.
//
// The scratch buffer (partial histograms + selection state) is a caller
// workspace of jq_loss_clip_workspace_bytes() bytes, or, for
// jq_loss_clip_noalloc, a static device array of the same size.

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/jaqmc_amd.h"
#include "kernels.h"

int jq_fail(int code, const std::string& msg);

namespace {

constexpr int kHistBlocks = 256;  // workgroups of a histogram pass (at most)
constexpr int kHistThreads = 256;
constexpr int kSel = 4;           // selection slots (two quantiles x floor/ceil)

// Selection state of one order statistic, in the scratch buffer.
struct Sel {
  unsigned long long prefix, mask;
  long long k;  // rank still to find below the current prefix
  long long n;  // non-NaN count
  double value;
  double pad[3];
};
// Clip state: the two quantile values, the window and the means.
struct ClipState {
  double q[2];
  double lo, hi;
  double mean_x, mean_c;
  double n;
  double pad;
};
// Scratch layout: [kSel Sel][ClipState][kHistBlocks x 256 uint32 partials]
constexpr size_t kScratchBytes = kSel * sizeof(Sel) + sizeof(ClipState) + (size_t)kHistBlocks * 256 * 4;
static_assert(kScratchBytes == JQ_LOSS_CLIP_SCRATCH_BYTES, "scratch size of the header");

__device__ inline unsigned long long key_of(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double val_of(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __builtin_bit_cast(double, b);
}

// Value fed to a selection: x, or |x - center| (center = a quantile result).
struct Src {
  const double* x;
  int B;
  const double* center;  // null: plain x
};
__device__ inline double src_val(const Src& s, int i) {
  const double v = s.x[i];
  return s.center ? fabs(v - *s.center) : v;
}

__global__ __launch_bounds__(kHistThreads) void hist_kernel(Src s, const Sel* __restrict__ sel, int shift,
                                                            unsigned* __restrict__ part) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long prefix = sel->prefix, mask = sel->mask;
  for (int i = blockIdx.x * kHistThreads + threadIdx.x; i < s.B; i += gridDim.x * kHistThreads) {
    const double v = src_val(s, i);
    if (isnan(v)) continue;
    const unsigned long long k = key_of(v);
    if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & 255u], 1u);
  }
  __syncthreads();
  part[blockIdx.x * 256 + threadIdx.x] = h[threadIdx.x];
}

// Sum the partials, pick the digit of the target rank.  On the first pass the
// rank is set from the non-NaN count: floor (side 0) or ceil (side 1) of q (n-1).
__global__ __launch_bounds__(256) void pick_kernel(Sel* __restrict__ sel, const unsigned* __restrict__ part,
                                                   int nblk, int shift, double q, int side) {
  __shared__ unsigned long long cum[257];
  unsigned long long t = 0;
  for (int b = 0; b < nblk; ++b) t += part[b * 256 + threadIdx.x];
  cum[threadIdx.x + 1] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    cum[0] = 0;
    for (int d = 0; d < 256; ++d) cum[d + 1] += cum[d];
    if (shift == 56) {
      sel->n = (long long)cum[256];
      const double pos = q * (double)(sel->n - 1);
      sel->k = sel->n > 0 ? (long long)(side ? ceil(pos) : floor(pos)) : 0;
    }
    const long long k = sel->k;
    int d = 0;
    while (d < 255 && (long long)cum[d + 1] <= k) ++d;
    sel->k = k - (long long)cum[d];
    sel->prefix |= (unsigned long long)d << shift;
    sel->mask |= 255ull << shift;
    if (shift == 0) sel->value = sel->n > 0 ? val_of(sel->prefix) : __builtin_nan("");
  }
}

__global__ void sel_init_kernel(Sel* sel) {
  if (threadIdx.x < kSel) {
    sel[threadIdx.x].prefix = 0;
    sel[threadIdx.x].mask = 0;
    sel[threadIdx.x].k = 0;
    sel[threadIdx.x].n = 0;
  }
}

// quantile q from the floor / ceil order statistics of slots a, b, weighted
// as jnp.quantile's linear method: lo (1 - f) + hi f with f = pos - floor(pos)
// (at q = 0.5 this equals nanmedian's midpoint (lo + hi) / 2 exactly).
__global__ void quantile_kernel(const Sel* sel, int a, int b, double q, double* out) {
  const long long n = sel[a].n;
  const double pos = q * (double)(n - 1);
  const double f = pos - floor(pos);
  const double lo = sel[a].value, hi = sel[b].value;
  *out = n > 0 ? lo * (1.0 - f) + hi * f : __builtin_nan("");
}

__global__ void window_kernel(ClipState* st, int method, double scale) {
  if (method == 0) {  // mad: q[0] median, q[1] median absolute deviation
    st->lo = st->q[0] - scale * st->q[1];
    st->hi = st->q[0] + scale * st->q[1];
  } else if (method == 1) {  // iqr: q[0] q1, q[1] q3
    const double iqr = st->q[1] - st->q[0];
    st->lo = st->q[0] - scale * iqr;
    st->hi = st->q[1] + scale * iqr;
  } else {
    st->lo = -__builtin_inf();
    st->hi = __builtin_inf();
  }
}

// clipped values, then nanmeans of x and of the clipped values (one workgroup)
__global__ __launch_bounds__(1024) void clip_mean_kernel(const double* __restrict__ x, int B, ClipState* st,
                                                         double* __restrict__ clipped) {
  __shared__ double s1[16], s2[16], s3[16];
  const double lo = st->lo, hi = st->hi;
  double a = 0.0, c = 0.0, n = 0.0;
  for (int i = threadIdx.x; i < B; i += blockDim.x) {
    const double v = x[i];
    const double cv = v < lo ? lo : (v > hi ? hi : v);  // NaN stays NaN (jnp.clip)
    if (clipped) clipped[i] = cv;
    if (!isnan(v)) {
      a += v;
      c += cv;
      n += 1.0;
    }
  }
  a = wave_sum_d(a);
  c = wave_sum_d(c);
  n = wave_sum_d(n);
  if ((threadIdx.x & 63) == 0) {
    s1[threadIdx.x >> 6] = a;
    s2[threadIdx.x >> 6] = c;
    s3[threadIdx.x >> 6] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double A = 0, C = 0, Nn = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      A += s1[i];
      C += s2[i];
      Nn += s3[i];
    }
    st->mean_x = A / Nn;
    st->mean_c = C / Nn;
    st->n = Nn;
  }
}

__global__ void weights_kernel(const double* __restrict__ x, int B, const ClipState* st,
                               double* __restrict__ w, double* __restrict__ stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double lo = st->lo, hi = st->hi, mc = st->mean_c, n = st->n;
  if (i < B) {
    const double v = x[i];
    const double cv = v < lo ? lo : (v > hi ? hi : v);
    w[i] = isnan(cv) ? 0.0 : 2.0 * (cv - mc) / n;
  }
  if (i == 0 && stats) {
    stats[0] = st->mean_x;
    stats[1] = st->mean_c;
    stats[2] = lo;
    stats[3] = hi;
    stats[4] = n;
  }
}

int select_into(const Src& src, Sel* sel, int slot, double q, int side, unsigned* part, hipStream_t s) {
  const int nblk = std::min(kHistBlocks, std::max(1, (src.B + kHistThreads * 4 - 1) / (kHistThreads * 4)));
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(hist_kernel, dim3(nblk), dim3(kHistThreads), 0, s, src, sel + slot, shift, part);
    hipLaunchKernelGGL(pick_kernel, dim3(1), dim3(256), 0, s, sel + slot, part, nblk, shift, q, side);
  }
  return 0;
}

int loss_clip(const double* x, int B, int method, double scale, double* clipped, double* weights, double* stats,
              void* scratch, hipStream_t s) {
  Sel* sel = (Sel*)scratch;
  ClipState* st = (ClipState*)(sel + kSel);
  unsigned* part = (unsigned*)(st + 1);
  hipLaunchKernelGGL(sel_init_kernel, dim3(1), dim3(64), 0, s, sel);
  const Src plain{x, B, nullptr};
  if (method == 0) {
    select_into(plain, sel, 0, 0.5, 0, part, s);
    select_into(plain, sel, 1, 0.5, 1, part, s);
    hipLaunchKernelGGL(quantile_kernel, dim3(1), dim3(1), 0, s, sel, 0, 1, 0.5, &st->q[0]);
    const Src dev{x, B, &st->q[0]};
    select_into(dev, sel, 2, 0.5, 0, part, s);
    select_into(dev, sel, 3, 0.5, 1, part, s);
    hipLaunchKernelGGL(quantile_kernel, dim3(1), dim3(1), 0, s, sel, 2, 3, 0.5, &st->q[1]);
  } else if (method == 1) {
    select_into(plain, sel, 0, 0.25, 0, part, s);
    select_into(plain, sel, 1, 0.25, 1, part, s);
    select_into(plain, sel, 2, 0.75, 0, part, s);
    select_into(plain, sel, 3, 0.75, 1, part, s);
    hipLaunchKernelGGL(quantile_kernel, dim3(1), dim3(1), 0, s, sel, 0, 1, 0.25, &st->q[0]);
    hipLaunchKernelGGL(quantile_kernel, dim3(1), dim3(1), 0, s, sel, 2, 3, 0.75, &st->q[1]);
  }
  hipLaunchKernelGGL(window_kernel, dim3(1), dim3(1), 0, s, st, method, scale);
  hipLaunchKernelGGL(clip_mean_kernel, dim3(1), dim3(1024), 0, s, x, B, st, clipped);
  hipLaunchKernelGGL(weights_kernel, dim3((B + 255) / 256), dim3(256), 0, s, x, B, st, weights, stats);
  return 0;
}

// The static scratch of jq_loss_clip_noalloc: one copy per device, held until
// the process ends.  Host code never uses the array's name as an address; it
// uses the device address from hipGetSymbolAddress, looked up once per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static void* scratch_address() {
  static std::mutex mu;
  static void* ptr[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    ptr[dev] = p;
  }
  return ptr[dev];
}

int check_args(const double* x, int B, int method, double* weights) {
  if (B < 1 || B > JQ_LOSS_CLIP_MAX_B || !x || !weights || method < 0 || method > 2)
    return jq_fail(JQ_EINVAL, "jq_loss_clip: bad argument (1 <= B <= JQ_LOSS_CLIP_MAX_B, method 0 mad / 1 iqr / 2 none)");
  return 0;
}

}  // namespace

extern "C" {

size_t jq_loss_clip_workspace_bytes(int B) {
  (void)B;
  return kScratchBytes;
}

int jq_loss_clip(const double* x, int B, int method, double scale, double* clipped, double* weights, double* stats,
                 void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_args(x, B, method, weights);
  if (rc) return rc;
  if (!workspace || ws_bytes < kScratchBytes)
    return jq_fail(JQ_EINVAL, "jq_loss_clip: workspace smaller than jq_loss_clip_workspace_bytes()");
  loss_clip(x, B, method, scale, clipped, weights, stats, workspace, (hipStream_t)stream);
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_loss_clip");
  return JQ_OK;
}

int jq_loss_clip_noalloc(const double* x, int B, int method, double scale, double* clipped, double* weights,
                         double* stats, void* stream) {
  int rc = check_args(x, B, method, weights);
  if (rc) return rc;
  if (B > JQ_LOSS_CLIP_NOALLOC_MAX_B)
    return jq_fail(JQ_EINVAL, "jq_loss_clip_noalloc: B exceeds JQ_LOSS_CLIP_NOALLOC_MAX_B");
  void* p = scratch_address();
  if (!p) return jq_fail(JQ_EHIP, "jq_loss_clip_noalloc: device address of the static scratch not found");
  loss_clip(x, B, method, scale, clipped, weights, stats, p, (hipStream_t)stream);
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_loss_clip_noalloc");
  return JQ_OK;
}

}  // extern "C"
