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

namespace {

// ---- host-side weight packing ----------------------------------------------
std::vector<uint16_t> to_bf16_bits(const std::vector<float>& v, std::vector<uint16_t>* lo) {
  std::vector<uint16_t> hi(v.size());
  if (lo) lo->resize(v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    // round to nearest even
    uint32_t b;
    std::memcpy(&b, &v[i], 4);
    uint32_t r = b + 0x7FFFu + ((b >> 16) & 1u);
    if ((b & 0x7F800000u) == 0x7F800000u) r = b;  // inf/nan passthrough
    hi[i] = (uint16_t)(r >> 16);
    if (lo) {
      uint32_t hb = (uint32_t)hi[i] << 16;
      float hf;
      std::memcpy(&hf, &hb, 4);
      float rem = v[i] - hf;
      uint32_t c;
      std::memcpy(&c, &rem, 4);
      uint32_t rr = c + 0x7FFFu + ((c >> 16) & 1u);
      (*lo)[i] = (uint16_t)(rr >> 16);
    }
  }
  return hi;
}

}  // namespace

int jq_model::upload(const void* host, size_t bytes, void** dev) {
  if (hipMalloc(dev, bytes ? bytes : 16) != hipSuccess) return jq_fail(JQ_ENOMEM, "hipMalloc failed for weights");
  allocs.push_back(*dev);
  ws_bytes += bytes;
  if (bytes && hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice) != hipSuccess)
    return jq_fail(JQ_EHIP, "hipMemcpy of weights failed");
  return 0;
}

// Upload a GEMM weight stored transposed (ncols x Kpad, zero padded), in the
// element format of the model's precision.
int jq_model::upload_wt(const std::vector<float>& wt, void** hi, void** lo) {
  if (prec == JQ_PREC_FP32) {
    *lo = nullptr;
    return upload(wt.data(), wt.size() * 4, hi);
  }
  std::vector<uint16_t> l;
  std::vector<uint16_t> h = to_bf16_bits(wt, prec == JQ_PREC_BF16X3 ? &l : nullptr);
  int rc = upload(h.data(), h.size() * 2, hi);
  if (rc) return rc;
  if (prec == JQ_PREC_BF16X3) return upload(l.data(), l.size() * 2, lo);
  *lo = nullptr;
  return 0;
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
  if (!m || !elec || !logpsi || B < 0) return jq_fail(JQ_EINVAL, "jq_logpsi: bad argument");
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
  if (!m || !elec || !e_kin || !e_pot || !e_loc || B < 0) return jq_fail(JQ_EINVAL, "jq_local_energy: bad argument");
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
  if (!m || !elec || !logp || !n_accept || B < 0 || n_steps < 0) return jq_fail(JQ_EINVAL, "jq_mh_steps: bad argument");
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
  if (!x || !out || B < 0) return jq_fail(JQ_EINVAL, "jq_reduce_mean_var: bad argument");
  launch_reduce_mean_var(x, B, out, (hipStream_t)stream);
  return JQ_OK;
}

int jq_reduce_accepts(const int32_t* n_accept, int B, double* out, void* stream) {
  if (!n_accept || !out || B < 0) return jq_fail(JQ_EINVAL, "jq_reduce_accepts: bad argument");
  launch_reduce_accepts(n_accept, B, out, (hipStream_t)stream);
  return JQ_OK;
}

}  // extern "C"
