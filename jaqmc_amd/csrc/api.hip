// C ABI of the MI355X VMC hot path: model creation (weight packing), the
// forward / forward-Laplacian pipelines over walker chunks, the Metropolis
// loop and the FP64 reductions.  See include/jaqmc_amd.h for the contract.

#include <algorithm>
#include <cmath>
#include <cstring>
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


int jq_model::upload(const void* host, size_t bytes, void** dev) {
  if (hipMalloc(dev, bytes ? bytes : 16) != hipSuccess) return jq_fail(JQ_ENOMEM, "hipMalloc failed for weights");
  allocs.push_back(*dev);
  ws_bytes += bytes;
  if (bytes && hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice) != hipSuccess)
    return jq_fail(JQ_EHIP, "hipMemcpy of weights failed");
  return 0;
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
int jq_model::upload_wt(const std::vector<float>& wt, void** hi, void** lo) {
  *lo = nullptr;
  *hi = nullptr;
  if (prec == JQ_PREC_FP32) return upload(wt.data(), wt.size() * 4, hi);
  const int ns = prec == JQ_PREC_BF16X6 ? 3 : (prec == JQ_PREC_BF16X3 ? 2 : 1);
  std::vector<uint16_t> planes(wt.size() * ns);
  for (size_t i = 0; i < wt.size(); ++i) {
    float r = wt[i];
    for (int s = 0; s < ns; ++s) {
      const uint16_t h = bf16_rne(r);
      planes[s * wt.size() + i] = h;
      r -= bf16_to_f(h);
    }
  }
  return upload(planes.data(), planes.size() * 2, lo);
}

void* jq_model::alloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
  allocs.push_back(p);
  ws_bytes += bytes;
  return p;
}

jq_model::~jq_model() {
  for (void* p : allocs) (void)hipFree(p);
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
  (void)hipDeviceSynchronize();
  delete m;
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
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_logpsi");
  return JQ_OK;
}

int jq_local_energy(jq_model* m, const float* elec, int B, double* e_kin, double* e_pot, double* e_loc,
                    float* logpsi, float* grad, void* stream) {
  if (!m || B < 0 || (B > 0 && (!elec || !e_kin || !e_pot || !e_loc)))
    return jq_fail(JQ_EINVAL, "jq_local_energy: bad argument");
  if (B == 0) return JQ_OK;
  hipStream_t s = (hipStream_t)stream;
  float* lp = logpsi;
  if (!lp) {
    int rc = m->ensure_scratch(B);
    if (rc) return rc;
    lp = m->lp;
  }
  for (int b0 = 0; b0 < B; b0 += m->chunk) {
    const int Bc = std::min(m->chunk, B - b0);
    int rc = m->forward(elec, b0, Bc, true, lp, nullptr, e_kin, e_pot, e_loc, grad, s);
    if (rc) return rc;
  }
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_local_energy");
  return JQ_OK;
}

int jq_mh_steps(jq_model* m, float* elec, float* logp, int B, int n_steps, float width, const float* normals,
                const float* uniforms, uint64_t seed, uint64_t ctr, int32_t* n_accept, void* stream) {
  if (!m || B < 0 || n_steps < 0 || (B > 0 && (!elec || !logp || !n_accept)))
    return jq_fail(JQ_EINVAL, "jq_mh_steps: bad argument");
  if (B == 0 || n_steps == 0) return JQ_OK;
  hipStream_t s = (hipStream_t)stream;
  int rc = m->ensure_scratch(B);
  if (rc) return rc;
  const long n3 = 3L * m->N;
  for (int st = 0; st < n_steps; ++st) {
    launch_propose(elec, m->prop, B, m->N, width, normals ? normals + (long)st * B * n3 : nullptr, seed, ctr + st, s);
    for (int b0 = 0; b0 < B; b0 += m->chunk_fwd) {
      const int Bc = std::min(m->chunk_fwd, B - b0);
      rc = m->forward(m->prop, b0, Bc, false, m->lp, nullptr, nullptr, nullptr, nullptr, nullptr, s);
      if (rc) return rc;
    }
    launch_accept(elec, m->prop, logp, m->lp, B, m->N, uniforms ? uniforms + (long)st * B : nullptr, seed, ctr + st,
                  n_accept, s);
  }
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_mh_steps");
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
