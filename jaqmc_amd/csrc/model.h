// Device-resident model: packed weights, walker-chunk workspace, pipelines.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "../../include/jaqmc_amd.h"
#include "kernels.h"

int jq_fail(int code, const std::string& msg);
// Per-device stream for model set-up and teardown (api.hip).
hipStream_t jq_setup_stream();

// Sequential reader over the flat FP32 parameter pack.
struct PackReader {
  const float* p;
  size_t n, pos = 0;
  bool ok = true;
  const float* take(size_t k) {
    if (pos + k > n) {
      ok = false;
      return nullptr;
    }
    const float* r = p + pos;
    pos += k;
    return r;
  }
};

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
  // parameter-gradient pass: FP32 weights in their original (in, out)
  // orientation, padded to K = d1 columns, and pack offsets
  float* Wm_o = nullptr;                 // (kmain, d1): rows [h | pair means]
  float* Wg_o = nullptr;                 // (kg, d1): rows [spin means]
  long off_w = 0, off_b = 0, off_w2 = 0, off_b2 = 0;
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
  // parameter-gradient pass: FP32 weights in their original (in, out)
  // orientation (the q columns carrying 1/sqrt(dh)), output widths padded to
  // 32 for the adjoint GEMMs, and pack offsets
  float* Wqkv_o = nullptr;                   // (D, 3D)
  float* Wo_o = nullptr;                     // (D, D)
  std::vector<float*> Wm_o;                  // (in, round_up(out, 32))
  long off_ln1s = 0, off_ln1b = 0, off_wq = 0, off_bq = 0, off_wk = 0, off_bk = 0, off_wv = 0, off_bv = 0;
  long off_wo = 0, off_bo = 0, off_ln2s = 0, off_ln2b = 0;
  std::vector<long> off_mw, off_mb;
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

  // output layers (with orbitals_spin_split=False both channel slots point at
  // the one shared kernel / envelope)
  int korb = 0;  // padded K of the orbital GEMM
  int norb = 1;  // orbital kernels in the pack: nch, or 1 when shared
  void* Worb[2] = {nullptr, nullptr};
  void* Worb_lo[2] = {nullptr, nullptr};
  float* borb[2] = {nullptr, nullptr};  // orbital bias (bias_orbitals) or null
  float* pi[2] = {nullptr, nullptr};
  float* sigma[2] = {nullptr, nullptr};
  int ln_mode = JQ_LN_PRE;  // Psiformer LayerNorm placement
  int options = 0;          // JQ_OPT_* bits

  // parameter-gradient pass (FermiNet, created with JQ_OPT_WITH_GRADIENTS)
  bool grads = false;
  int gchunk = 0;           // walkers per gradient chunk
  long n_params = 0;        // floats in the parameter pack
  long off_orb[2] = {0, 0}, off_env[2] = {0, 0};
  float* Worb_o[2] = {nullptr, nullptr};  // (d_last, KNpad) original orientation
  long ldko = 0;                          // KN padded to 32
  std::vector<void*> Xt;                  // tape: layer inputs X_0 .. X_L (value rows)
  std::vector<float*> H2t;                // tape: pair stream h2_0 .. h2_{L-1}
  std::vector<void*> gt;                  // tape: spin means g_0 .. g_{L-1}
  float *dH[2] = {nullptr, nullptr}, *dZ = nullptr, *dX = nullptr, *dG = nullptr, *dg = nullptr;
  float *dH2[2] = {nullptr, nullptr}, *dZ2 = nullptr, *dPhi = nullptr, *dEnv = nullptr, *Pw = nullptr;
  long pw_cap = 0, ldd = 0;
  double* gacc = nullptr;                 // FP64 gradient accumulator (n_params)
  int alloc_grad_workspace(size_t tape_budget);
  int forward_ferminet_tape(const float* elec_chunk, int Bc, hipStream_t s);
  int vjp(const float* elec, int B, const double* cot, float* grad, hipStream_t s);
  // Psiformer: offsets of the input layer, orbital biases and Jastrow; the
  // tape H_0 .. H_L (residual stream), per layer the attention input XIN
  // (LN1 output pre-LN), head-split Q/K/V, attention output, H1 (after the
  // attention block), the MLP input M0 (LN2 output pre-LN), the MLP outputs
  // and, post-LN, the sums U1, U2 the LayerNorms normalise; the features F0
  long off_win = 0, off_bin = 0, off_borb[2] = {0, 0}, off_jas = 0;
  float* F0t = nullptr;
  std::vector<float*> Ht, XINt, QKVt, Ot, H1t, M0t, U1t, U2t;
  std::vector<std::vector<float*>> MHt;
  float *dH1 = nullptr, *dM[2] = {nullptr, nullptr}, *dO = nullptr, *dQKV = nullptr, *dXin = nullptr;
  float *Pln = nullptr, *dPhiM = nullptr, *jtmp = nullptr;
  int alloc_grad_workspace_psi(size_t tape_budget);
  int forward_psiformer_tape(const float* elec_chunk, int Bc, hipStream_t s);
  int vjp_psiformer(const float* elec, int B, const double* cot, hipStream_t s);

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
  // Metropolis scratch, allocated at create: proposals of one forward chunk
  // and log|psi| of one chunk
  float* prop = nullptr;
  float* lp = nullptr;
  bool skip_det = false;  // orbitals export: stop after Phi and the envelope
  // streams this model has been used on, each with an event recorded after the
  // model's last call there (destroy waits for these, nothing device-wide)
  std::vector<std::pair<hipStream_t, hipEvent_t>> used;
  void touch(hipStream_t s);

  ~jq_model();
  int init(const jq_model_desc& desc, const float* packed, size_t n_floats);
  int init_ferminet(const float* packed, size_t n_floats);
  int init_psiformer(const float* packed, size_t n_floats);
  int alloc_mh_scratch();
  // Run one walker chunk [b0, b0 + Bc) of `elec` (global walker indexing) and
  // write per-walker outputs at offset b0.
  // logpsi is indexed by global walker (b0 + w), or by chunk walker w when
  // lp_local (the model's own scratch)
  int forward(const float* elec, int b0, int Bc, bool lap, float* logpsi, float* sign, double* e_kin,
              double* e_pot, double* e_loc, float* grad, hipStream_t s, bool lp_local = false);
  int forward_ferminet(const float* elec_chunk, int Bc, bool lap, hipStream_t s);
  int forward_psiformer(const float* elec_chunk, int Bc, bool lap, hipStream_t s);
  // orbital projection of the last residual stream x, envelope, determinants
  int psi_outputs(const float* elec_chunk, const float* x, int Bc, bool lap, hipStream_t s);
  int orbitals_and_dets(const float* elec_chunk, const void* Xlast, int Bc, bool lap, hipStream_t s);

  int read_outputs(PackReader& rd, int d_last);
  int upload(const void* host, size_t bytes, void** dev);
  int upload_wt(const std::vector<float>& wt, int ncols, void** hi, void** lo);
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
