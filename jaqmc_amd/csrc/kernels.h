// Host-side launchers of the gfx950 kernels (one declaration per kernel family).
#pragma once

#include "common.h"

enum { PREC_F32 = 0, PREC_BF16 = 1, PREC_BF16X3 = 2, PREC_BF16X6 = 3, PREC_FP16X3 = 4 };
enum { ACT_NONE = 0, ACT_TANH = 1 };

// Y[rows, ncols] = epilogue(X[rows, K] . Wt[ncols, K]^T)
//
// Operands are K-contiguous ("NT") FP32: X rows and Wt rows (one per output
// column).  K must be a multiple of 32 and ldx, ldw multiples of 4; padding
// columns of X and Wt must be zero.  In Laplacian mode (R > 1) rows come in
// units of R = Rp tangent rows (value, Laplacian, Jacobian...) and one
// workgroup owns one whole unit, so the elementwise forward-Laplacian rule of
// the activation can be applied in the epilogue:
//   y   = acc + G[walker(unit), r] + bias * [r == value]
//   tanh: value -> t, Jacobian -> (1-t^2) J, Laplacian -> (1-t^2) L - 2t(1-t^2) sum_c J_c^2
//   residual: out = (out + Res[row]) * res_scale
struct GemmArgs {
  const void* X; long ldx;
  const void* Wt; long ldw;        // FP32 (ncols, K) transposed weight (PREC_F32)
  const void* Wt_lo; long wplane;  // split modes: NS BF16 planes of the split weight,
                                   // plane s at Wt_lo + s * wplane (elements).  PREC_FP16X3
                                   // weights hold three BF16 planes (used by the value-only
                                   // launches), then two FP16 planes of the column-scaled
                                   // weight and the ncols inverse column scales (FP32)
  void* Y; long ldy;
  int rows, K, ncols;
  const float* bias;   // (ncols) or null
  const float* G; long ldg;  // per-walker broadcast term or null
  int R;               // rows per unit (Rp) in Laplacian mode, 1 otherwise
  int act;             // ACT_*
  const void* Res; long ldres;  // residual source (same element type as Y) or null
  float res_scale;
  int Bc, n_up, n_dn;  // unit -> walker map (spin-blocked units)
  int wmaj;            // units are walker-major (unit = w*N + i) instead of spin-blocked
  int nbatch;          // > 1: blockIdx.z batches with X += z*sxb, Y += z*syb (W shared)
  long sxb, syb;
  // head-split output (Psiformer QKV): column n = part*D + h*dh + d of unit
  // (w, i), row r goes to Y[(((w*3 + part)*H + h)*N + i)*R*dh + r*dh + d]
  int hs_H, hs_dh, hs_D, hs_N;
  int linear_tiles;  // 1: plain row-tile-major workgroup order (no XCD grouping)
  int lrp;           // rows are tangent rows of units of lrp rows (value, Laplacian, Jacobian...);
                     // 0: value rows only.  PREC_FP16X3 scales the value and Laplacian rows
                     // of each unit by powers of two before the FP16 split (set by the caller
                     // for R = 1 launches on Laplacian-mode data; R > 1 implies lrp = R)
  int zpad;          // > ncols: columns [ncols, zpad) of Y are written as zeros
  int rgrp;          // > 0 (plain rows only): row r is row r % rgrp of group r / rgrp, at
  long sxg, syg;     // X + (r / rgrp) * sxg + (r % rgrp) * ldx and Y likewise with syg, ldy
  // LayerNorm of the output in Laplacian mode (the forward-Laplacian rule of
  // layernorm_lap_kernel): Yln (ld ncols) = LN(Y) with scale ln_s, bias ln_b,
  // rows [ln_R, R) of every unit zero.  Only where one workgroup holds whole
  // units and every column (gemm_ln_fusable); null: none.
  float* Yln;
  const float* ln_s;
  const float* ln_b;
  float ln_eps;
  int ln_R;
};
// Whether a Laplacian-mode GEMM of R rows per unit and ncols output columns
// can produce the LayerNorm of its output in its epilogue (GemmArgs::Yln).
bool gemm_ln_fusable(int R, int ncols, int prec);

// Batched GEMM on activation operands (both FP32 in HBM, split on load for the
// BF16 modes): for z = w*bh + h, Y_z (+)= X_z . op(W_z) where op(W) is W^T
// (Wt given as ncols x K, "NT") or W (given as K x ncols, "NN").  Rows of W
// beyond kvalid read as zero.  Output columns may be remapped in segments:
// col n -> (n / ycw) * ycs + n % ycw (ycw = 0: identity).
struct BGemmArgs {
  const float* X; long ldx; long sxw, sxh;
  const float* W; long ldw; long sww, swh;
  float* Y; long ldy; long syw, syh;
  const float* Res;  // accumulate: Y = X.op(W) + Res (same addressing as Y) or null
  int rows, K, ncols, kvalid;
  int nw, nh;        // batch grid
  int nn;            // 0: NT, 1: NN
  int ycw, ycs;
};
int launch_bgemm(const BGemmArgs& a, int prec, hipStream_t s);

// prec selects the MFMA arithmetic (PREC_*); all operands and results are
// FP32 in memory (y_bf16 must be false).
int launch_gemm(const GemmArgs& a, int prec, bool y_bf16, hipStream_t s);

// ---- FermiNet feature / stream kernels (ferminet.hip) ----
struct FermiDims {
  int Bc, N, n_up, n_dn, M, K;  // walkers in chunk, electrons, atoms, dets
  int Rp;                       // rows per unit (1 in forward mode)
  int nch;                      // non-empty spin channels (1 or 2)
  int wmaj;                     // walker-major units (Psiformer) instead of spin-blocked
};
__host__ __device__ inline int unit_idx(const FermiDims& d, int w, int i) {
  return d.wmaj ? w * d.N + i : unit_of(w, i, d.Bc, d.n_up, d.n_dn);
}
// Floats per envelope element in env[walker, i, k*N + o, :]: value,
// Laplacian and gradient in Laplacian mode (Rp > 1), the value alone in
// forward mode.
__host__ __device__ inline int env_stride(const FermiDims& d) { return d.Rp > 1 ? 5 : 1; }
// Layer-0 one-electron input (ae features) into X[:, :, 0:4M] and the pair
// stream input H2 (RP2 rows per pair: 8 in Laplacian mode, 1 in forward mode).
int launch_ferminet_features(const FermiDims& d, const float* elec, const float* atoms,
                             void* X, long ldx, bool x_bf16, float* H2, hipStream_t s);
// One pair-stream layer: H2out = tanh(H2in W + b) (+ residual).
int launch_pair_layer(const FermiDims& d, const float* H2in, int fin, float* H2out, int fout,
                      const float* W, const float* b, bool residual, hipStream_t s);
// Means over the first electron index of the pair stream, expanded to dense
// tangent rows, written into X[:, :, col0 : col0 + nch*f2].
int launch_pair_means(const FermiDims& d, const float* H2, int f2, void* X, long ldx, int col0,
                      bool x_bf16, hipStream_t s);
// Spin-channel means of X[:, :, 0:f1] -> g[walker, r, ch*f1 + f]  (ldg row stride).
int launch_spin_means(const FermiDims& d, const void* X, long ldx, int f1, void* g, long ldg,
                      bool bf16io, hipStream_t s);
// Envelope value / Laplacian / gradient: env[walker, i, k*N + o, env_stride(d)].
int launch_envelope(const FermiDims& d, const float* elec, const float* atoms, const float* const* pi,
                    const float* const* sigma, int env_kind, float* env, hipStream_t s);

// ---- determinants (det.hip) ----
// Per (walker, det): FP64 Gauss-Jordan inverse of A = Phi * env, log|det|, sign,
// and (Laplacian mode) d logdet / dx_c and the Laplacian of log|det|.
// out layout per (w, k): [logdet, sign, lap, grad[3N]] doubles (stride 3 + 3N).
int launch_det(const FermiDims& d, const float* Phi, long ldphi, const float* env, double* out,
               hipStream_t s);
// Combine determinants: log|psi| (+ Jastrow for Psiformer), E_kin, V, E_L.
struct CombineArgs {
  const double* det; int B_off;  // chunk offset into the caller's arrays
  const float* elec; const float* atoms; const float* charges;
  int jastrow; float alpha_par, alpha_anti;
  float* logpsi; float* sign; double* e_kin; double* e_pot; double* e_loc; float* grad;
  int lapmode;
  int lp_local;  // logpsi indexed by chunk walker (model scratch) instead of B_off + walker
};
int launch_combine(const FermiDims& d, const CombineArgs& a, hipStream_t s);

// ---- Metropolis (mcmc.hip) ----
// Walkers [w0, w0 + B) of a batch: elec/prop/normals/uniforms/logp/n_accept
// point at walker w0; the Philox streams are keyed by the global index w0 + w.
// width_dev (device, the sampler state's stddev as FP64) overrides width when set.
int launch_propose(const float* elec, float* prop, int B, int N, float width, const double* width_dev,
                   const float* normals, uint64_t seed, uint64_t ctr, int w0, hipStream_t s);
int launch_accept(float* elec, const float* prop, float* logp, const float* logpsi_prop, int B, int N,
                  const float* uniforms, uint64_t seed, uint64_t ctr, int w0, int32_t* n_accept, hipStream_t s);
// MCMCSampler.step's adaptation on the device state (see jq_sample).
int launch_mcmc_adapt(double* state, const int32_t* n_accept, int B, int steps, int af, double lo, double hi,
                      double* pmove_out, hipStream_t s);
// A[w][k][i][o] = Phi value row * envelope value (orbitals export).
int launch_orbital_export(const FermiDims& d, const float* Phi, long ldphi, const float* env, float* out,
                          hipStream_t s);
int launch_reduce_mean_var(const double* x, int B, double* out, hipStream_t s);
int launch_reduce_accepts(const int32_t* x, int B, double* out, hipStream_t s);
int launch_scale_f32(float* x, int n, float a, hipStream_t s);

// ---- Psiformer kernels (psiformer.hip) ----
// Layer-0 atom-electron features (rescaled [log(1+r), d log(1+r)/r] when
// rescale, else [r, d]) + spin column into X[:, :, 0:kin).
int launch_psi_features(const FermiDims& d, const float* elec, const float* atoms, float* X, int kin, int rescale,
                        hipStream_t s);
// LayerNorm forward Laplacian over the feature axis, one unit (Rp rows) per
// workgroup: y = (x - mean) rsqrt(var + eps) * scale + bias.
int launch_layernorm_lap(int units, int Rp, int R, int D, const float* x, float* y, const float* scale,
                         const float* bias, float eps, hipStream_t s);
// Softmax forward Laplacian of attention scores for every (walker, head, query).
struct SoftmaxArgs {
  const float* S1; const float* S2; const float* S3; int lds;  // score partials
  float* A1; int lda1;   // [z][(i,c)][j] rows c < Rp, columns lda1
  float* A2;             // [z][i][(j,c)] = 2 * Jacobian rows (Laplacian mode)
  int N, Rp, R, nz;
};
int launch_softmax_lap(const SoftmaxArgs& a, hipStream_t s);

// Fused forward-Laplacian attention (attention.hip): QKV head-split
// [w][part][h][i][Rp][dh]; output row (i, c) of walker w, head h at
// O + ((w*N + i)*Rp + c)*ldo + h*dh.  nz = walkers * heads.
struct AttnArgs {
  const float* QKV;
  float* O; int ldo;
  int N, Rp, R, H, dh, nz;
  int h3;  // 1: the q.k products in the FP16X3 split (PREC_FP16X3), else FP32
};
bool attn_lap_supported(int N, int dh);
int launch_attn_lap(const AttnArgs& a, hipStream_t s);
// Value-only attention of the forward pass (Rp = 1): O[w][i][h*dh + d].
bool attn_fwd_supported(int N, int dh);
int launch_attn_fwd(const AttnArgs& a, hipStream_t s);
