// Device-resident model: packed weights, walker-chunk workspace, pipelines.
#pragma once

#include <string>
#include <vector>

#include "../../include/jaqmc_amd.h"
#include "kernels.h"

int jq_fail(int code, const std::string& msg);

// One FermiNet layer (FermiLayers, wavefunction/backbone/ferminet.py:31-63).
// Envelope parameters arrive as (orbital o, atom I, det k) (the Flax
// (N_orb, M, K) layout); the device keeps them as [I][k][o] so that the
// threads of an envelope workgroup, which walk (k, o), read them coalesced.
inline std::vector<float> envelope_device_layout(const float* p, int N, int M, int K) {
  std::vector<float> t((size_t)N * M * K);
  for (int o = 0; o < N; ++o)
    for (int I = 0; I < M; ++I)
      for (int k = 0; k < K; ++k) t[((size_t)I * K + k) * N + o] = p[((size_t)o * M + I) * K + k];
  return t;
}

struct FermiLayer {
  int f1, f2;        // one- / two-electron input widths
  int d1, d2;        // output widths
  int kmain;         // K of the per-electron GEMM: [h (f1) | pair means (nch*f2)] padded to 32
  int kg;            // K of the spin-mean GEMM: nch*f1 padded to 32
  bool has_pair;     // two-electron stream updated in this layer (l < L-1)
  bool res1, res2;   // residual (x + y)/sqrt2 applies (shapes match)
  void *Wm = nullptr, *Wm_lo = nullptr;  // (d1, kmain) transposed, GEMM element format
  void *Wg = nullptr, *Wg_lo = nullptr;  // (d1, kg)
  float* b = nullptr;                    // (d1)
  float* W2 = nullptr;                   // (f2, d2) row-major FP32 (pair stream)
  float* b2 = nullptr;
};

// One Psiformer layer (PsiformerLayer, wavefunction/backbone/psiformer.py:64-98).
struct PsiLayer {
  float *ln1_s, *ln1_b, *ln2_s, *ln2_b;
  void *Wqkv = nullptr, *Wqkv_lo = nullptr;  // (3D, D) transposed: q, k, v stacked
  float* bqkv = nullptr;                     // (3D)
  void *Wo = nullptr, *Wo_lo = nullptr;      // (D, D)
  float* bo = nullptr;
  std::vector<void*> Wm, Wm_lo;              // MLP (out, in_pad) transposed
  std::vector<float*> bm;
  std::vector<int> mdims;                    // D, hidden..., D
};

struct jq_model {
  jq_model_desc d{};
  int kind = 0;
  int N = 0, T = 0, M = 0, K = 0, nch = 1, L = 0, prec = 0;
  int n_up = 0, n_dn = 0;
  int Rp = 16;       // tangent rows per unit (3N + 2 padded to a GEMM unit height)
  bool xbf = false;  // activations stored as BF16
  float* atoms = nullptr;
  float* charges = nullptr;
  std::vector<float> atoms_h, charges_h;
  float jas[2] = {1.f, 1.f};

  // FermiNet
  std::vector<FermiLayer> fl;
  // Psiformer
  std::vector<PsiLayer> pl;
  int D = 0, H = 0, dh = 0, kin = 0;  // model dim, heads, head dim, padded input K
  void *Win = nullptr, *Win_lo = nullptr;
  float* bin = nullptr;

  // output layers
  int korb = 0;  // padded K of the orbital GEMM
  void* Worb[2] = {nullptr, nullptr};
  void* Worb_lo[2] = {nullptr, nullptr};
  float* pi[2] = {nullptr, nullptr};
  float* sigma[2] = {nullptr, nullptr};

  // workspace (sized for `chunk` walkers in Laplacian mode; forward mode reuses
  // it for `chunk_fwd` walkers)
  int chunk = 0, chunk_fwd = 0;
  size_t ws_bytes = 0;
  std::vector<void*> allocs;
  void* X[2] = {nullptr, nullptr};
  void* g = nullptr;
  float* G = nullptr;
  float* H2[2] = {nullptr, nullptr};
  float* Phi = nullptr;
  long ldphi = 0;
  float* env = nullptr;
  double* det = nullptr;
  // Psiformer scratch
  void* P[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  float* stats = nullptr;
  // Metropolis scratch (grown on demand)
  float* prop = nullptr;
  float* lp = nullptr;
  int prop_cap = 0;

  ~jq_model();
  int init(const jq_model_desc& desc, const float* packed, size_t n_floats);
  int init_ferminet(const float* packed, size_t n_floats);
  int init_psiformer(const float* packed, size_t n_floats);
  int ensure_scratch(int B);
  // Run one walker chunk [b0, b0 + Bc) of `elec` (global walker indexing) and
  // write per-walker outputs at offset b0.
  int forward(const float* elec, int b0, int Bc, bool lap, float* logpsi, float* sign, double* e_kin,
              double* e_pot, double* e_loc, float* grad, hipStream_t s);
  int forward_ferminet(const float* elec_chunk, int Bc, bool lap, hipStream_t s);
  int forward_psiformer(const float* elec_chunk, int Bc, bool lap, hipStream_t s);
  int orbitals_and_dets(const float* elec_chunk, const void* Xlast, int Bc, bool lap, hipStream_t s);

  int upload(const void* host, size_t bytes, void** dev);
  int upload_wt(const std::vector<float>& wt, void** hi, void** lo);
  void* alloc(size_t bytes);
  size_t es() const { return xbf ? 2 : 4; }
  FermiDims dims(int Bc, bool lap) const {
    FermiDims fd;
    fd.Bc = Bc;
    fd.N = N;
    fd.n_up = n_up;
    fd.n_dn = n_dn;
    fd.M = M;
    fd.K = K;
    fd.Rp = lap ? Rp : 1;
    fd.nch = nch;
    fd.wmaj = kind == JQ_KIND_PSIFORMER ? 1 : 0;
    return fd;
  }
};
