// C ABI of the MI355X VMC hot path: model creation (weight packing), the
// forward / forward-Laplacian pipelines over walker chunks, the Metropolis
// loop and the FP64 reductions.  See include/jaqmc_amd.h for the contract.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jaqmc_amd.h"
#include "kernels.h"
#include "model.h"

thread_local std::string g_jq_err;

int jq_fail(int code, const std::string& msg) {
  g_jq_err = msg;
  return code;
}

int gemm_unit_rows(int R);

// Per-device stream for model set-up and teardown: device memory comes from
// the stream-ordered allocator on it, so a model's destroy waits only for the
// streams that model ran on (through their events) and never synchronises the
// device.  Created once per device and kept for the life of the process.
hipStream_t jq_setup_stream() {
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!streams[dev]) {
    if (hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess) {
      streams[dev] = nullptr;
      return nullptr;
    }
    // keep up to 64 GiB of released model memory in the device's pool, so the
    // next model's workspace is carved from it instead of the driver
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t keep = 64ull << 30;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
  }
  return streams[dev];
}

int jq_model::upload(const void* host, size_t bytes, void** dev) {
  *dev = alloc(bytes);
  if (!*dev) return jq_fail(JQ_ENOMEM, "device allocation failed for weights");
  hipStream_t st = jq_setup_stream();
  if (bytes && (hipMemcpyAsync(*dev, host, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess))
    return jq_fail(JQ_EHIP, "copy of weights to the device failed");
  return 0;
}

void jq_model::touch(hipStream_t s) {
  for (auto& u : used)
    if (u.first == s) {
      (void)hipEventRecord(u.second, s);
      return;
    }
  hipEvent_t e;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return;
  used.emplace_back(s, e);
  (void)hipEventRecord(e, s);
}

namespace {
uint16_t bf16_rne(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  if ((b & 0x7F800000u) == 0x7F800000u) return (uint16_t)(b >> 16);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
float bf16_to_f(uint16_t h) {
  const uint32_t b = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
}  // namespace

// Upload a GEMM weight stored transposed (ncols x Kpad, zero padded): FP32 for
// PREC_F32, otherwise the NS-term BF16 split (x = t0 + t1 + ..., each term the
// round-to-nearest-even BF16 of the remaining residual) as NS planes.
// PREC_FP16X3 uploads three BF16 planes (value-only launches run the BF16X6
// kernel), then two FP16 planes of the column-scaled weight (column n times
// 2^s_n, s_n chosen so its largest entry lies in [2^13, 2^14) and the low
// terms stay normal), then the ncols inverse scales 2^-s_n and the ncols
// scales 2^s_n as FP32.
int jq_model::upload_wt(const std::vector<float>& wt, int ncols, void** hi, void** lo) {
  *lo = nullptr;
  *hi = nullptr;
  if (prec == JQ_PREC_FP32) return upload(wt.data(), wt.size() * 4, hi);
  const int ns = (prec == JQ_PREC_BF16X6 || prec == JQ_PREC_FP16X3) ? 3 : (prec == JQ_PREC_BF16X3 ? 2 : 1);
  const size_t n = wt.size();
  const size_t extra = prec == JQ_PREC_FP16X3 ? 2 * n + 4 * (size_t)ncols : 0;  // in uint16 units
  std::vector<uint16_t> planes(n * ns + extra);
  for (size_t i = 0; i < n; ++i) {
    float r = wt[i];
    for (int s = 0; s < ns; ++s) {
      const uint16_t h = bf16_rne(r);
      planes[s * n + i] = h;
      r -= bf16_to_f(h);
    }
  }
  if (prec == JQ_PREC_FP16X3) {
    const size_t kpad = n / (size_t)ncols;
    _Float16* f16p = reinterpret_cast<_Float16*>(planes.data() + 3 * n);
    std::vector<float> sc(2 * (size_t)ncols);
    for (int c = 0; c < ncols; ++c) {
      float mx = 0.f;
      for (size_t k = 0; k < kpad; ++k) mx = std::max(mx, std::fabs(wt[c * kpad + k]));
      int sh = 0;
      if (mx > 0.f && std::isfinite(mx)) {
        int e;
        (void)std::frexp(mx, &e);
        sh = std::min(60, std::max(-60, 14 - e));
      }
      const float s_c = std::ldexp(1.f, sh);
      sc[c] = std::ldexp(1.f, -sh);
      sc[ncols + c] = s_c;
      for (size_t k = 0; k < kpad; ++k) {
        const float x = wt[c * kpad + k] * s_c;
        const _Float16 h = (_Float16)x;
        f16p[c * kpad + k] = h;
        f16p[n + c * kpad + k] = (_Float16)(x - (float)h);
      }
    }
    std::memcpy(planes.data() + 5 * n, sc.data(), sc.size() * 4);
  }
  return upload(planes.data(), planes.size() * 2, lo);
}

void* jq_model::alloc(size_t bytes) {
  void* p = nullptr;
  hipStream_t st = jq_setup_stream();
  if (hipMallocAsync(&p, bytes ? bytes : 16, st) != hipSuccess) return nullptr;
  if (hipStreamSynchronize(st) != hipSuccess) {
    (void)hipFreeAsync(p, st);
    return nullptr;
  }
  allocs.push_back(p);
  ws_bytes += bytes;
  return p;
}

// Stream-ordered release: the set-up stream first waits for the last call of
// this model on every stream it used.
jq_model::~jq_model() {
  if (allocs.empty() && used.empty()) return;  // rejected before any device work
  hipStream_t st = jq_setup_stream();
  for (auto& u : used) (void)hipStreamWaitEvent(st, u.second, 0);
  for (void* p : allocs) (void)hipFreeAsync(p, st);
  for (auto& u : used) (void)hipEventDestroy(u.second);
}

extern "C" {

const char* jq_last_error(void) { return g_jq_err.c_str(); }

int jq_model_create(const jq_model_desc* desc, const float* packed, size_t n_floats, jq_model** out) {
  if (!desc || !packed || !out) return jq_fail(JQ_EINVAL, "null argument");
  *out = nullptr;
  jq_model* m = new jq_model();
  int rc = m->init(*desc, packed, n_floats);
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return JQ_OK;
}

int jq_model_destroy(jq_model* m) {
  if (!m) return JQ_OK;
  delete m;  // no host or device-wide synchronisation (see ~jq_model)
  return JQ_OK;
}

size_t jq_model_workspace_bytes(const jq_model* m) { return m ? m->ws_bytes : 0; }

int jq_logpsi(jq_model* m, const float* elec, int B, float* logpsi, float* sign, void* stream) {
  if (!m || B < 0 || (B > 0 && (!elec || !logpsi))) return jq_fail(JQ_EINVAL, "jq_logpsi: bad argument");
  if (B == 0) return JQ_OK;
  hipStream_t s = (hipStream_t)stream;
  for (int b0 = 0; b0 < B; b0 += m->chunk_fwd) {
    const int Bc = std::min(m->chunk_fwd, B - b0);
    int rc = m->forward(elec, b0, Bc, false, logpsi, sign, nullptr, nullptr, nullptr, nullptr, s);
    if (rc) return rc;
  }
  m->touch(s);
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_logpsi");
  return JQ_OK;
}

int jq_orbitals(jq_model* m, const float* elec, int B, float* orbitals, void* stream) {
  if (!m || B < 0 || (B > 0 && (!elec || !orbitals))) return jq_fail(JQ_EINVAL, "jq_orbitals: bad argument");
  if (B == 0) return JQ_OK;
  hipStream_t s = (hipStream_t)stream;
  m->skip_det = true;
  int rc = 0;
  for (int b0 = 0; b0 < B && !rc; b0 += m->chunk_fwd) {
    const int Bc = std::min(m->chunk_fwd, B - b0);
    rc = m->forward(elec, b0, Bc, false, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, s);
    if (!rc)
      launch_orbital_export(m->dims(Bc, false), m->Phi, m->ldphi, m->env,
                            orbitals + (long)b0 * m->K * m->N * m->N, s);
  }
  m->skip_det = false;
  if (rc) return rc;
  m->touch(s);
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_orbitals");
  return JQ_OK;
}

int jq_local_energy(jq_model* m, const float* elec, int B, double* e_kin, double* e_pot, double* e_loc,
                    float* logpsi, float* grad, void* stream) {
  if (!m || B < 0 || (B > 0 && (!elec || !e_kin || !e_pot || !e_loc)))
    return jq_fail(JQ_EINVAL, "jq_local_energy: bad argument");
  if (B == 0) return JQ_OK;
  hipStream_t s = (hipStream_t)stream;
  // without a caller buffer, log|psi| goes to the model's one-chunk scratch
  float* lp = logpsi ? logpsi : m->lp;
  for (int b0 = 0; b0 < B; b0 += m->chunk) {
    const int Bc = std::min(m->chunk, B - b0);
    int rc = m->forward(elec, b0, Bc, true, lp, nullptr, e_kin, e_pot, e_loc, grad, s, logpsi == nullptr);
    if (rc) return rc;
  }
  m->touch(s);
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_local_energy");
  return JQ_OK;
}

int jq_mh_steps(jq_model* m, float* elec, float* logp, int B, int n_steps, float width, const float* normals,
                const float* uniforms, uint64_t seed, uint64_t ctr, int32_t* n_accept, void* stream) {
  if (!m || B < 0 || n_steps < 0 || (B > 0 && (!elec || !logp || !n_accept)))
    return jq_fail(JQ_EINVAL, "jq_mh_steps: bad argument");
  if (B == 0 || n_steps == 0) return JQ_OK;
  hipStream_t s = (hipStream_t)stream;
  const long n3 = 3L * m->N;
  // walker chunks outermost: each chunk makes all its moves through the
  // model's one-chunk proposal scratch (walkers are independent and the
  // Philox streams are keyed by the global walker index)
  for (int b0 = 0; b0 < B; b0 += m->chunk_fwd) {
    const int Bc = std::min(m->chunk_fwd, B - b0);
    float* ec = elec + (long)b0 * n3;
    for (int st = 0; st < n_steps; ++st) {
      launch_propose(ec, m->prop, Bc, m->N, width, nullptr, normals ? normals + ((long)st * B + b0) * n3 : nullptr,
                     seed, ctr + st, b0, s);
      int rc = m->forward(m->prop, 0, Bc, false, m->lp, nullptr, nullptr, nullptr, nullptr, nullptr, s, true);
      if (rc) return rc;
      launch_accept(ec, m->prop, logp + b0, m->lp, Bc, m->N, uniforms ? uniforms + (long)st * B + b0 : nullptr, seed,
                    ctr + st, b0, n_accept + b0, s);
    }
  }
  m->touch(s);
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_mh_steps");
  return JQ_OK;
}

int jq_sample(jq_model* m, float* elec, float* logp, int B, int n_outer, int steps, double* state, int adapt_frequency,
              double pmove_lo, double pmove_hi, uint64_t seed, uint64_t ctr, int32_t* n_accept, double* pmove_out,
              void* stream) {
  if (!m || B < 0 || n_outer < 0 || steps < 0 || adapt_frequency < 1 ||
      (B > 0 && (!elec || !logp || !n_accept || !state)))
    return jq_fail(JQ_EINVAL, "jq_sample: bad argument");
  if (B == 0 || n_outer == 0 || steps == 0) return JQ_OK;  // MCMCSampler.step with steps == 0 is a no-op
  hipStream_t s = (hipStream_t)stream;
  const long n3 = 3L * m->N;
  for (int o = 0; o < n_outer; ++o) {
    for (int b0 = 0; b0 < B; b0 += m->chunk_fwd) {
      const int Bc = std::min(m->chunk_fwd, B - b0);
      float* ec = elec + (long)b0 * n3;
      // logp = 2 log|psi| at the current walkers, recomputed each sampler step
      int rc = m->forward(elec, b0, Bc, false, logp, nullptr, nullptr, nullptr, nullptr, nullptr, s);
      if (rc) return rc;
      launch_scale_f32(logp + b0, Bc, 2.f, s);
      for (int t = 0; t < steps; ++t) {
        const uint64_t c = ctr + (uint64_t)o * steps + t;
        launch_propose(ec, m->prop, Bc, m->N, 0.f, state, nullptr, seed, c, b0, s);
        rc = m->forward(m->prop, 0, Bc, false, m->lp, nullptr, nullptr, nullptr, nullptr, nullptr, s, true);
        if (rc) return rc;
        launch_accept(ec, m->prop, logp + b0, m->lp, Bc, m->N, nullptr, seed, c, b0, n_accept + b0, s);
      }
    }
    launch_mcmc_adapt(state, n_accept, B, steps, adapt_frequency, pmove_lo, pmove_hi,
                      pmove_out ? pmove_out + o : nullptr, s);
  }
  m->touch(s);
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_sample");
  return JQ_OK;
}

int jq_reduce_mean_var(const double* x, int B, double* out, void* stream) {
  if (!out || B < 0 || (B > 0 && !x)) return jq_fail(JQ_EINVAL, "jq_reduce_mean_var: bad argument");
  launch_reduce_mean_var(x, B, out, (hipStream_t)stream);
  return JQ_OK;
}

int jq_reduce_accepts(const int32_t* n_accept, int B, double* out, void* stream) {
  if (!out || B < 0 || (B > 0 && !n_accept)) return jq_fail(JQ_EINVAL, "jq_reduce_accepts: bad argument");
  launch_reduce_accepts(n_accept, B, out, (hipStream_t)stream);
  return JQ_OK;
}

}  // extern "C"
