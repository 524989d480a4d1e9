// FermiNet parameter gradient: forward pass with a per-layer tape, then the
// reverse sweep (kernels in grad.hip, dense products on the GEMM family in
// FP32).  See grad.hip for the derivative rules and their reference sources.

#include <algorithm>
#include <cmath>

#include "model.h"
#include "prof.h"

int launch_det_vjp(const FermiDims& d, const float* Phi, long ldphi, const float* env, const double* det,
                   const double* cot, int b0, float* dPhi, long lddphi, float* denv, hipStream_t s);
int launch_env_grad(const FermiDims& d, const float* elec, const float* atoms, const float* pi0, const float* pi1,
                    const float* sg0, const float* sg1, int kind, int nsets, const float* denv, double* acc,
                    long off0, long off1, hipStream_t s);
int launch_tanh_bwd(long rows, int n, const float* hnew, long ldn, const float* hold, long ldo, const float* dhnew,
                    long lddn, int res, float* dz, long lddz, float* dhold, long lddo, hipStream_t s);
int launch_walker_sum(const FermiDims& d, const float* dz, long lddz, int n, float* dG, long lddg, hipStream_t s);
int launch_one_adj(const FermiDims& d, int f1, const float* dX, long lddx, const float* dg, long lddg, float* dh,
                   long lddh, hipStream_t s);
int launch_pair_adj(const FermiDims& d, int f2, const float* dX, long lddx, int col0, float* dh2, hipStream_t s);
int launch_pair_dx(long npairs, int fin, int fout, const float* dz2, const float* W2, int res, const float* dh2new,
                   float* dh2old, hipStream_t s);
int launch_wgrad(long rows, const float* X, long ldx, int Kx, const float* dY, long ldy, int Ny, float* P, long pcap,
                 double* acc, long off, int row0, int ld, hipStream_t s);
int launch_fill_f64(double* p, long n, double v, hipStream_t s);
int launch_fill_f32(float* p, long n, float v, hipStream_t s);
int launch_f64_to_f32(const double* a, long n, float* out, hipStream_t s);

int jq_model::alloc_grad_workspace(size_t budget) {
  const long KN = (long)K * N;
  long dmax = 32, kmax = 32, kgm = 32, f2m = 4, d2m = 4;
  size_t tape = 0;
  for (int l = 0; l < L; ++l) {
    const FermiLayer& ly = fl[l];
    dmax = std::max(dmax, (long)ly.d1);
    kmax = std::max(kmax, (long)ly.kmain);
    kgm = std::max(kgm, (long)ly.kg);
    f2m = std::max(f2m, (long)ly.f2);
    if (ly.has_pair) d2m = std::max(d2m, (long)ly.d2);
    tape += (size_t)N * ly.kmain + (size_t)N * N * ly.f2 + ly.kg;
  }
  tape += (size_t)N * korb;
  ldd = dmax;
  const size_t back = (size_t)N * dmax * 3 + (size_t)N * kmax + dmax + kgm + (size_t)N * N * (2 * f2m + d2m) +
                      (size_t)N * ldko + (size_t)N * KN;
  const size_t per_walker = 4 * (tape + back) + 64;
  gchunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)chunk_fwd, budget / per_walker));
  const long G = gchunk;
  Xt.resize(L + 1);
  H2t.resize(L);
  gt.resize(L);
  for (int l = 0; l <= L; ++l) {
    const long ld = l < L ? fl[l].kmain : korb;
    Xt[l] = alloc((size_t)G * N * ld * 4);
    if (!Xt[l]) return jq_fail(JQ_ENOMEM, "gradient tape allocation failed");
  }
  for (int l = 0; l < L; ++l) {
    H2t[l] = (float*)alloc((size_t)G * N * N * fl[l].f2 * 4);
    gt[l] = alloc((size_t)G * fl[l].kg * 4);
    if (!H2t[l] || !gt[l]) return jq_fail(JQ_ENOMEM, "gradient tape allocation failed");
  }
  for (int i = 0; i < 2; ++i) {
    dH[i] = (float*)alloc((size_t)G * N * dmax * 4);
    dH2[i] = (float*)alloc((size_t)G * N * N * f2m * 4);
  }
  dZ = (float*)alloc((size_t)G * N * dmax * 4);
  dX = (float*)alloc((size_t)G * N * kmax * 4);
  dG = (float*)alloc((size_t)G * dmax * 4);
  dg = (float*)alloc((size_t)G * kgm * 4);
  dZ2 = (float*)alloc((size_t)G * N * N * d2m * 4);
  dPhi = (float*)alloc((size_t)G * N * ldko * 4);
  dEnv = (float*)alloc((size_t)G * N * KN * 4);
  pw_cap = 16l << 20;
  Pw = (float*)alloc((size_t)pw_cap * 4);
  gacc = (double*)alloc((size_t)n_params * 8);
  if (!dH[0] || !dH[1] || !dH2[0] || !dH2[1] || !dZ || !dX || !dG || !dg || !dZ2 || !dPhi || !dEnv || !Pw || !gacc)
    return jq_fail(JQ_ENOMEM, "gradient workspace allocation failed");
  // padding columns of dPhi (KN .. ldko) are read by the orbital adjoint GEMM
  // and must stay zero: the determinant adjoint writes columns < KN only
  hipStream_t st = jq_setup_stream();
  launch_fill_f32(dPhi, (long)G * N * ldko, 0.f, st);
  if (hipStreamSynchronize(st) != hipSuccess) return jq_fail(JQ_EHIP, "gradient workspace initialisation failed");
  return 0;
}

// Value-mode FermiNet forward with every layer's input kept: X_l, h2_l, g_l.
int jq_model::forward_ferminet_tape(const float* elec, int Bc, hipStream_t s) {
  const FermiDims fd = dims(Bc, false);
  int rc = launch_ferminet_features(fd, elec, atoms, Xt[0], fl[0].kmain, false, H2t[0], s);
  for (int l = 0; l < L && !rc; ++l) {
    const FermiLayer& ly = fl[l];
    void* Xc = Xt[l];
    void* Xn = Xt[l + 1];
    rc = launch_pair_means(fd, H2t[l], ly.f2, Xc, ly.kmain, ly.f1, false, s);
    if (!rc) rc = launch_spin_means(fd, Xc, ly.kmain, ly.f1, gt[l], ly.kg, false, s);
    if (rc) break;
    GemmArgs ga{};
    ga.X = gt[l];
    ga.ldx = ly.kg;
    ga.Wt = ly.Wg;
    ga.Wt_lo = ly.Wg_lo;
    ga.wplane = (long)ly.d1 * ly.kg;
    ga.ldw = ly.kg;
    ga.Y = G;
    ga.ldy = ly.d1;
    ga.rows = Bc;
    ga.K = ly.kg;
    ga.ncols = ly.d1;
    ga.R = 1;
    ga.act = ACT_NONE;
    if (launch_gemm(ga, prec, false, s)) return jq_fail(JQ_EINVAL, "spin-mean GEMM launch rejected");
    GemmArgs gm{};
    gm.X = Xc;
    gm.ldx = ly.kmain;
    gm.Wt = ly.Wm;
    gm.Wt_lo = ly.Wm_lo;
    gm.wplane = (long)ly.d1 * ly.kmain;
    gm.ldw = ly.kmain;
    gm.Y = Xn;
    gm.ldy = (l + 1 < L) ? fl[l + 1].kmain : korb;
    gm.rows = Bc * N;
    gm.K = ly.kmain;
    gm.ncols = ly.d1;
    gm.bias = ly.b;
    gm.G = G;
    gm.ldg = ly.d1;
    gm.R = 1;
    gm.act = ACT_TANH;
    gm.Res = ly.res1 ? Xc : nullptr;
    gm.ldres = ly.kmain;
    gm.res_scale = 0.70710678118654752f;
    gm.Bc = Bc;
    gm.n_up = n_up;
    gm.n_dn = n_dn;
    if (launch_gemm(gm, prec, false, s)) return jq_fail(JQ_EINVAL, "layer GEMM launch rejected");
    if (ly.has_pair) rc = launch_pair_layer(fd, H2t[l], ly.f2, H2t[l + 1], ly.d2, ly.W2, ly.b2, ly.res2, s);
  }
  if (rc) return jq_fail(JQ_EHIP, "FermiNet kernel launch failed");
  return orbitals_and_dets(elec, Xt[L], Bc, false, s);
}

// dY = dZ W^T on the FP32 GEMM: X = dz (rows x K), Wt = W in (out, K) order.
static int gemm_nt(const float* Xp, long ldx, const float* W, int K_, int ncols, float* Y, long ldy, long rows,
                   hipStream_t s) {
  GemmArgs g{};
  g.X = Xp;
  g.ldx = ldx;
  g.Wt = W;
  g.ldw = K_;
  g.Y = Y;
  g.ldy = ldy;
  g.rows = (int)rows;
  g.K = K_;
  g.ncols = ncols;
  g.R = 1;
  g.act = ACT_NONE;
  g.res_scale = 1.f;
  return launch_gemm(g, PREC_F32, false, s);
}

int jq_model::vjp(const float* elec, int B, const double* cot, float* grad, hipStream_t s) {
  launch_fill_f64(gacc, n_params, 0.0, s);
  const long KN = (long)K * N;
  for (int b0 = 0; b0 < B; b0 += gchunk) {
    const int Bc = std::min(gchunk, B - b0);
    const float* ec = elec + (long)b0 * T;
    const FermiDims fd = dims(Bc, false);
    int rc = forward_ferminet_tape(ec, Bc, s);
    if (rc) return rc;
    if (launch_det_vjp(fd, Phi, ldphi, env, det, cot, b0, dPhi, ldko, dEnv, s))
      return jq_fail(JQ_EINVAL, "determinant adjoint supports N <= 64");
    if (d.envelope != JQ_ENV_NULL)
      launch_env_grad(fd, ec, atoms, pi[0], pi[1], sigma[0], sigma[1], d.envelope, norb, dEnv, gacc, off_env[0],
                      off_env[1], s);
    // orbital projection: dW_orb (channel rows, spin-blocked units), dh_L = dPhi W_orb^T
    const int dL = fl[L - 1].d1;
    float* dh = dH[0];
    float* dho = dH[1];
    for (int c = 0; c < nch; ++c) {
      const long u0 = (nch == 1 || c == 0) ? 0 : (long)Bc * n_up;
      const long nu = nch == 1 ? (long)Bc * N : (c == 0 ? (long)Bc * n_up : (long)Bc * n_dn);
      const int set = norb == 2 ? c : 0;
      rc = launch_wgrad(nu, (const float*)Xt[L] + u0 * korb, korb, dL, dPhi + u0 * ldko, ldko, (int)KN, Pw, pw_cap,
                        gacc, off_orb[set], 0, (int)KN, s);
      rc |= gemm_nt(dPhi + u0 * ldko, ldko, Worb_o[set], (int)ldko, dL, dh + u0 * ldd, ldd, nu, s);
      if (rc) return jq_fail(JQ_EINVAL, "orbital adjoint launch rejected");
    }
    // layers, last to first
    const long units = (long)Bc * N, npairs = (long)Bc * N * N;
    float* dh2n = dH2[0];
    float* dh2o = dH2[1];
    for (int l = L - 1; l >= 0; --l) {
      const FermiLayer& ly = fl[l];
      const long ldn = (l + 1 < L) ? fl[l + 1].kmain : korb;
      // pair stream layer l (h2_l -> h2_{l+1}) needs the complete adjoint of h2_{l+1}
      if (ly.has_pair) {
        launch_tanh_bwd(npairs, ly.d2, H2t[l + 1], ly.d2, H2t[l], ly.f2, dh2n, ly.d2, ly.res2, dZ2, ly.d2, nullptr, 0,
                        s);
        rc = launch_wgrad(npairs, H2t[l], ly.f2, ly.f2, dZ2, ly.d2, ly.d2, Pw, pw_cap, gacc, ly.off_w2, 0, ly.d2, s);
        rc |= launch_wgrad(npairs, nullptr, 0, 1, dZ2, ly.d2, ly.d2, Pw, pw_cap, gacc, ly.off_b2, 0, ly.d2, s);
        if (rc) return jq_fail(JQ_EINVAL, "pair-stream adjoint launch rejected");
        if (l > 0) launch_pair_dx(npairs, ly.f2, ly.d2, dZ2, ly.W2, ly.res2, dh2n, dh2o, s);
      } else if (l > 0) {
        launch_fill_f32(dh2o, npairs * ly.f2, 0.f, s);
      }
      // one-electron stream: dz, weight and bias gradients; the adjoint of h_l
      // starts from the residual branch (or zero, with its own width f1)
      if (l > 0 && !ly.res1) launch_fill_f32(dho, units * ldd, 0.f, s);
      launch_tanh_bwd(units, ly.d1, (const float*)Xt[l + 1], ldn, (const float*)Xt[l], ly.kmain, dh, ldd, ly.res1,
                      dZ, ldd, (l > 0 && ly.res1) ? dho : nullptr, ldd, s);
      const int in_rows_h = ly.f1, g_row0 = ly.f1, p_row0 = ly.f1 * (1 + nch);
      rc = launch_wgrad(units, (const float*)Xt[l], ly.kmain, in_rows_h, dZ, ldd, ly.d1, Pw, pw_cap, gacc, ly.off_w, 0,
                        ly.d1, s);
      rc |= launch_wgrad(units, (const float*)Xt[l] + ly.f1, ly.kmain, nch * ly.f2, dZ, ldd, ly.d1, Pw, pw_cap, gacc,
                         ly.off_w, p_row0, ly.d1, s);
      rc |= launch_wgrad(units, nullptr, 0, 1, dZ, ldd, ly.d1, Pw, pw_cap, gacc, ly.off_b, 0, ly.d1, s);
      launch_walker_sum(fd, dZ, ldd, ly.d1, dG, ly.d1, s);
      rc |= launch_wgrad(Bc, (const float*)gt[l], ly.kg, nch * ly.f1, dG, ly.d1, ly.d1, Pw, pw_cap, gacc, ly.off_w,
                         g_row0, ly.d1, s);
      if (rc) return jq_fail(JQ_EINVAL, "layer adjoint launch rejected");
      if (l > 0) {
        // adjoints of h_l (spin means and the layer input) and of h2_l (pair means)
        rc = gemm_nt(dG, ly.d1, ly.Wg_o, ly.d1, nch * ly.f1, dg, ly.kg, Bc, s);
        rc |= gemm_nt(dZ, ldd, ly.Wm_o, ly.d1, ly.f1 + nch * ly.f2, dX, ly.kmain, units, s);
        if (rc) return jq_fail(JQ_EINVAL, "input adjoint GEMM launch rejected");
        launch_one_adj(fd, ly.f1, dX, ly.kmain, dg, ly.kg, dho, ldd, s);
        launch_pair_adj(fd, ly.f2, dX, ly.kmain, ly.f1, dh2o, s);
        std::swap(dh, dho);
        std::swap(dh2n, dh2o);
      }
    }
  }
  launch_f64_to_f32(gacc, n_params, grad, s);
  return 0;
}

extern "C" {

int jq_logpsi_vjp(jq_model* m, const float* elec, int B, const double* cot, float* grad, void* stream) {
  if (!m || B < 0 || (B > 0 && (!elec || !cot)) || !grad) return jq_fail(JQ_EINVAL, "jq_logpsi_vjp: bad argument");
  if (!m->grads)
    return jq_fail(JQ_EINVAL, "jq_logpsi_vjp: the model was created without JQ_OPT_WITH_GRADIENTS");
  hipStream_t s = (hipStream_t)stream;
  int rc = m->vjp(elec, B, cot, grad, s);
  if (rc) return rc;
  m->touch(s);
  if (hipGetLastError() != hipSuccess) return jq_fail(JQ_EHIP, "kernel launch failed in jq_logpsi_vjp");
  return JQ_OK;
}

size_t jq_model_param_count(const jq_model* m) { return m ? (size_t)m->n_params : 0; }

}  // extern "C"
