// Parameter gradients (FermiNet, Psiformer): forward pass with a per-layer
// tape, then the reverse sweep (kernels in grad.hip, dense products on the GEMM family in
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
  if (kind == JQ_KIND_PSIFORMER) {
    const int rc = vjp_psiformer(elec, B, cot, s);
    if (rc) return rc;
    launch_f64_to_f32(gacc, n_params, grad, s);
    return 0;
  }
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

// ---------------------------------------------------------------------------
// Psiformer, value path:
//   H_0 = F0 Win + bin;  per layer (pre-LN; no LayerNorm drops LN1, LN2)
//   XIN = LN1(H),  q|k|v = XIN Wqkv + b,  O = MHA(q, k, v),  H1 = H + O Wo + bo,
//   M0 = LN2(H1),  m = tanh(... tanh(M0 W_0 + b_0) ... W_last + b_last),
//   H' = H1 + m;  post-LN:  U1 = H + O Wo + bo (XIN = H),  H1 = LN1(U1),
//   M0 = H1,  U2 = H1 + m,  H' = LN2(U2);
//   Phi = H_L W_orb (+ b_orb) per spin channel, then envelope and determinants.
// The reverse sweep runs the same sequence backwards; each dense layer's
// weight gradient is X^T dY over the chunk's rows (wgrad), its input adjoint
// dY W^T on the FP32 GEMM with the weights in their (in, out) orientation.
// ---------------------------------------------------------------------------
int launch_ln_bwd(long rows, int D, const float* x, long ldx, const float* s, const float* dy, long lddy, float eps,
                  const float* base, long ldb, float* out, long ldo, float* P, long ldp, hipStream_t st);
size_t attn_bwd_smem(int N, int dh);
int launch_attn_bwd(const float* QKV, const float* dO, long lddo, float* dQKV, long ldq, int Bc, int N, int H, int dh,
                    hipStream_t st);
int launch_jastrow_grad(const float* elec, int Bc, int N, int n_up, float ap, float aa, const double* cot, int b0,
                        float* out, hipStream_t st);
int launch_mask_rows(long units, int N, int i0, int i1, int ncols, const float* src, long lds, float* dst, long ldd,
                     hipStream_t st);
int launch_add_rows(long rows, int n, const float* a, long lda, const float* b, long ldb, float* out, long ldo,
                    hipStream_t st);
int launch_scale_f64(double* p, long n, double v, hipStream_t st);

int jq_model::alloc_grad_workspace_psi(size_t budget) {
  const long KN = (long)K * N;
  const bool pre = ln_mode == JQ_LN_PRE, post = ln_mode == JQ_LN_POST;
  long tape = kin + (long)(L + 1) * D, wmax = D;
  for (const PsiLayer& ly : pl) {
    tape += (pre || post ? 2L : 0L) * D + 3L * D + 2L * D;
    for (size_t j = 1; j < ly.mdims.size(); ++j) {
      tape += round_up(ly.mdims[j], 32);
      wmax = std::max(wmax, (long)round_up(ly.mdims[j], 32));
    }
  }
  const long back = 2L * D + 2L * wmax + wmax + D + 3L * D + D + D + 2L * ldko + KN;
  const size_t per_walker = 4 * (size_t)N * (tape + back) + 64;
  gchunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)chunk_fwd, budget / per_walker));
  const size_t U = (size_t)gchunk * N;
  auto fa = [&](size_t floats) { return (float*)alloc(std::max<size_t>(floats, 1) * 4); };
  F0t = fa(U * kin);
  Ht.assign(L + 1, nullptr);
  XINt.assign(L, nullptr);
  QKVt.assign(L, nullptr);
  Ot.assign(L, nullptr);
  H1t.assign(L, nullptr);
  M0t.assign(L, nullptr);
  U1t.assign(L, nullptr);
  U2t.assign(L, nullptr);
  MHt.assign(L, {});
  bool ok = F0t != nullptr;
  for (int l = 0; l <= L; ++l) ok = ok && (Ht[l] = fa(U * D)) != nullptr;
  for (int l = 0; l < L && ok; ++l) {
    XINt[l] = pre ? fa(U * D) : Ht[l];
    QKVt[l] = fa(U * 3 * D);
    Ot[l] = fa(U * D);
    H1t[l] = fa(U * D);
    M0t[l] = pre ? fa(U * D) : H1t[l];
    if (post) {
      U1t[l] = fa(U * D);
      U2t[l] = fa(U * D);
    }
    ok = XINt[l] && QKVt[l] && Ot[l] && H1t[l] && M0t[l] && (!post || (U1t[l] && U2t[l]));
    for (size_t j = 1; j < pl[l].mdims.size() && ok; ++j) {
      MHt[l].push_back(fa(U * round_up(pl[l].mdims[j], 32)));
      ok = MHt[l].back() != nullptr;
    }
  }
  dH[0] = fa(U * D);
  dH1 = fa(U * D);
  dM[0] = fa(U * wmax);
  dM[1] = fa(U * wmax);
  dZ = fa(U * wmax);
  dO = fa(U * D);
  dQKV = fa(U * 3 * D);
  dXin = fa(U * D);
  Pln = fa(U * D);
  dPhi = fa(U * ldko);
  dPhiM = fa(U * ldko);
  dEnv = fa(U * KN);
  jtmp = fa(2 * (size_t)gchunk);
  pw_cap = 16l << 20;
  Pw = fa(pw_cap);
  gacc = (double*)alloc((size_t)n_params * 8);
  if (!ok || !dH[0] || !dH1 || !dM[0] || !dM[1] || !dZ || !dO || !dQKV || !dXin || !Pln || !dPhi || !dPhiM ||
      !dEnv || !jtmp || !Pw || !gacc)
    return jq_fail(JQ_ENOMEM, "gradient workspace allocation failed");
  // padding columns of dPhi (KN .. ldko) are read by the orbital adjoint GEMM
  hipStream_t st = jq_setup_stream();
  launch_fill_f32(dPhi, (long)U * ldko, 0.f, st);
  if (hipStreamSynchronize(st) != hipSuccess) return jq_fail(JQ_EHIP, "gradient workspace initialisation failed");
  return 0;
}

// Value-mode Psiformer forward with the tape (the forward_psiformer sequence
// with every intermediate kept; FP32 attention on the VALU).
int jq_model::forward_psiformer_tape(const float* elec, int Bc, hipStream_t s) {
  const FermiDims fd = dims(Bc, false);
  const long units = (long)Bc * N;
  const bool pre = ln_mode == JQ_LN_PRE, post = ln_mode == JQ_LN_POST;
  auto gemm = [&](const float* X, long ldx, void* Wt, void* Wlo, long wplane, int K_, float* Y, long ldy, int ncols,
                  const float* bias, int act, const float* Res, int zpad) {
    GemmArgs g{};
    g.X = X;
    g.ldx = ldx;
    g.Wt = Wt;
    g.Wt_lo = Wlo;
    g.wplane = wplane;
    g.ldw = K_;
    g.Y = Y;
    g.ldy = ldy;
    g.rows = (int)units;
    g.K = K_;
    g.ncols = ncols;
    g.bias = bias;
    g.R = 1;
    g.act = act;
    g.Res = Res;
    g.ldres = ldy;
    g.res_scale = 1.f;
    g.Bc = Bc;
    g.n_up = n_up;
    g.n_dn = n_dn;
    g.wmaj = 1;
    g.zpad = zpad;
    return launch_gemm(g, prec, false, s);
  };
  int rc = launch_psi_features(fd, elec, atoms, F0t, kin, (options & JQ_OPT_NO_RESCALE) ? 0 : 1, s);
  rc |= gemm(F0t, kin, Win, Win_lo, (long)D * kin, kin, Ht[0], D, D, bin, ACT_NONE, nullptr, 0);
  for (int l = 0; l < L && !rc; ++l) {
    const PsiLayer& ly = pl[l];
    if (pre) rc |= launch_layernorm_lap((int)units, 1, 1, D, Ht[l], XINt[l], ly.ln1_s, ly.ln1_b, 1e-5f, s);
    GemmArgs g{};
    g.X = XINt[l];
    g.ldx = D;
    g.Wt = ly.Wqkv;
    g.Wt_lo = ly.Wqkv_lo;
    g.wplane = 3L * D * D;
    g.ldw = D;
    g.Y = QKVt[l];
    g.ldy = 3 * D;
    g.rows = (int)units;
    g.K = D;
    g.ncols = 3 * D;
    g.bias = ly.bqkv;
    g.R = 1;
    g.act = ACT_NONE;
    g.res_scale = 1.f;
    g.wmaj = 1;
    g.hs_H = H;
    g.hs_dh = dh;
    g.hs_D = D;
    g.hs_N = N;
    rc |= launch_gemm(g, prec, false, s);
    AttnArgs aa{};
    aa.QKV = QKVt[l];
    aa.O = Ot[l];
    aa.ldo = D;
    aa.N = N;
    aa.Rp = 1;
    aa.R = 1;
    aa.H = H;
    aa.dh = dh;
    aa.nz = Bc * H;
    rc |= launch_attn_fwd(aa, s);
    rc |= gemm(Ot[l], D, ly.Wo, ly.Wo_lo, (long)D * D, D, post ? U1t[l] : H1t[l], D, D, ly.bo, ACT_NONE, Ht[l], 0);
    if (post) rc |= launch_layernorm_lap((int)units, 1, 1, D, U1t[l], H1t[l], ly.ln1_s, ly.ln1_b, 1e-5f, s);
    if (pre) rc |= launch_layernorm_lap((int)units, 1, 1, D, H1t[l], M0t[l], ly.ln2_s, ly.ln2_b, 1e-5f, s);
    const float* cur = M0t[l];
    int curw = D;
    for (size_t j = 0; j < ly.Wm.size() && !rc; ++j) {
      const int outw = ly.mdims[j + 1], outp = round_up(outw, 32), ka = round_up(curw, 32);
      rc |= gemm(cur, j == 0 ? D : ka, ly.Wm[j], ly.Wm_lo[j], (long)outw * ka, ka, MHt[l][j], outp, outw, ly.bm[j],
                 ACT_TANH, nullptr, outp);
      cur = MHt[l][j];
      curw = outw;
    }
    if (!rc) launch_add_rows(units, D, H1t[l], D, MHt[l].back(), D, post ? U2t[l] : Ht[l + 1], D, s);
    if (post) rc |= launch_layernorm_lap((int)units, 1, 1, D, U2t[l], Ht[l + 1], ly.ln2_s, ly.ln2_b, 1e-5f, s);
  }
  if (rc) return jq_fail(JQ_EINVAL, "Psiformer tape launch rejected");
  return psi_outputs(elec, Ht[L], Bc, false, s);
}

// dY = dX W^T for the channel rows of every walker (walker-major row groups)
static int gemm_nt_groups(const float* Xp, long ldx, const float* W, int K_, int ncols, float* Y, long ldy, long rows,
                          int rgrp, long sxg, long syg, hipStream_t s) {
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
  g.rgrp = rgrp;
  g.sxg = sxg;
  g.syg = syg;
  return launch_gemm(g, PREC_F32, false, s);
}

int jq_model::vjp_psiformer(const float* elec, int B, const double* cot, hipStream_t s) {
  const long KN = (long)K * N;
  const bool pre = ln_mode == JQ_LN_PRE, post = ln_mode == JQ_LN_POST;
  auto wg = [&](long rows, const float* X, long ldx, int Kx, const float* dY, long ldy, int Ny, long off) {
    return launch_wgrad(rows, X, ldx, Kx, dY, ldy, Ny, Pw, pw_cap, gacc, off, 0, Ny, s);
  };
  for (int b0 = 0; b0 < B; b0 += gchunk) {
    const int Bc = std::min(gchunk, B - b0);
    const long units = (long)Bc * N;
    const float* ec = elec + (long)b0 * T;
    const FermiDims fd = dims(Bc, false);
    int rc = forward_psiformer_tape(ec, Bc, s);
    if (rc) return rc;
    if (launch_det_vjp(fd, Phi, ldphi, env, det, cot, b0, dPhi, ldko, dEnv, s))
      return jq_fail(JQ_EINVAL, "determinant adjoint supports N <= 64");
    if (d.envelope != JQ_ENV_NULL)
      launch_env_grad(fd, ec, atoms, pi[0], pi[1], sigma[0], sigma[1], d.envelope, norb, dEnv, gacc, off_env[0],
                      off_env[1], s);
    // orbital projection: weight (and bias) gradients from the channel's rows,
    // dH_L = dPhi W_orb^T per channel
    float* gH = dH[0];  // adjoint of the residual stream
    for (int c = 0; c < nch; ++c) {
      const int set = norb == 2 ? c : 0;
      const int i0 = (nch == 1 || c == 0) ? 0 : n_up, ns = nch == 1 ? N : (c == 0 ? n_up : n_dn);
      const float* dP = dPhi;
      if (nch == 2) {
        launch_mask_rows(units, N, i0, i0 + ns, (int)ldko, dPhi, ldko, dPhiM, ldko, s);
        dP = dPhiM;
      }
      rc = wg(units, Ht[L], D, D, dP, ldko, (int)KN, off_orb[set]);
      if (borb[c]) rc |= wg(units, nullptr, 0, 1, dP, ldko, (int)KN, off_borb[set]);
      rc |= gemm_nt_groups(dPhi + (long)i0 * ldko, ldko, Worb_o[set], (int)ldko, D, gH + (long)i0 * D, D,
                           (long)Bc * ns, ns, (long)N * ldko, (long)N * D, s);
      if (rc) return jq_fail(JQ_EINVAL, "orbital adjoint launch rejected");
    }
    for (int l = L - 1; l >= 0 && !rc; --l) {
      const PsiLayer& ly = pl[l];
      if (post) {  // H' = LN2(U2): the adjoint of U2 (= that of m) into dH1
        if (launch_ln_bwd(units, D, U2t[l], D, ly.ln2_s, gH, D, 1e-5f, nullptr, 0, dH1, D, Pln, D, s))
          return jq_fail(JQ_EINVAL, "LayerNorm adjoint supports D <= 512");
        rc = wg(units, nullptr, 0, 1, Pln, D, D, ly.off_ln2s);
        rc |= wg(units, nullptr, 0, 1, gH, D, D, ly.off_ln2b);
      }
      // MLP: dm starts as the adjoint of the MLP output (H' = H1 + m, or U2)
      const int nm = (int)ly.Wm.size();
      const float* dm = post ? dH1 : gH;
      long lddm = D;
      int mi = 0;
      for (int j = nm - 1; j >= 0 && !rc; --j) {
        const int inw = ly.mdims[j], outw = ly.mdims[j + 1], outp = round_up(outw, 32);
        if (outp != outw) launch_fill_f32(dZ, units * outp, 0.f, s);
        launch_tanh_bwd(units, outw, MHt[l][j], outp, nullptr, 0, dm, lddm, 0, dZ, outp, nullptr, 0, s);
        const float* xin = j == 0 ? M0t[l] : MHt[l][j - 1];
        const long ldxin = j == 0 ? D : round_up(inw, 32);
        rc = wg(units, xin, ldxin, inw, dZ, outp, outw, ly.off_mw[j]);
        rc |= wg(units, nullptr, 0, 1, dZ, outp, outw, ly.off_mb[j]);
        const long ldn = round_up(inw, 32);
        rc |= gemm_nt(dZ, outp, ly.Wm_o[j], outp, inw, dM[mi], ldn, units, s);
        dm = dM[mi];
        lddm = ldn;
        mi ^= 1;
      }
      if (rc) return jq_fail(JQ_EINVAL, "MLP adjoint launch rejected");
      // LN2 (pre-LN): dH1 = dH' + LN2'(dm); with no LayerNorm dH1 = dH' + dm;
      // post-LN: the adjoint of H1 is dU2 + dm, then H1 = LN1(U1) gives dU1 in dH1
      if (post) {
        launch_add_rows(units, D, dH1, D, dm, lddm, gH, D, s);
        launch_ln_bwd(units, D, U1t[l], D, ly.ln1_s, gH, D, 1e-5f, nullptr, 0, dH1, D, Pln, D, s);
        rc = wg(units, nullptr, 0, 1, Pln, D, D, ly.off_ln1s);
        rc |= wg(units, nullptr, 0, 1, gH, D, D, ly.off_ln1b);
      } else if (pre) {
        if (launch_ln_bwd(units, D, H1t[l], D, ly.ln2_s, dm, lddm, 1e-5f, gH, D, dH1, D, Pln, D, s))
          return jq_fail(JQ_EINVAL, "LayerNorm adjoint supports D <= 512");
        rc = wg(units, nullptr, 0, 1, Pln, D, D, ly.off_ln2s);
        rc |= wg(units, nullptr, 0, 1, dm, lddm, D, ly.off_ln2b);
      } else {
        launch_add_rows(units, D, gH, D, dm, lddm, dH1, D, s);
      }
      // attention output projection and the attention itself
      rc |= wg(units, Ot[l], D, D, dH1, D, D, ly.off_wo);
      rc |= wg(units, nullptr, 0, 1, dH1, D, D, ly.off_bo);
      rc |= gemm_nt(dH1, D, ly.Wo_o, D, D, dO, D, units, s);
      if (rc) return jq_fail(JQ_EINVAL, "attention projection adjoint launch rejected");
      if (launch_attn_bwd(QKVt[l], dO, D, dQKV, 3L * D, Bc, N, H, dh, s))
        return jq_fail(JQ_EINVAL, "attention adjoint supports N <= 64 and heads_dim <= 64");
      const long offw[3] = {ly.off_wq, ly.off_wk, ly.off_wv}, offb[3] = {ly.off_bq, ly.off_bk, ly.off_bv};
      for (int p = 0; p < 3; ++p) {
        rc |= wg(units, XINt[l], D, D, dQKV + (long)p * D, 3L * D, D, offw[p]);
        rc |= wg(units, nullptr, 0, 1, dQKV + (long)p * D, 3L * D, D, offb[p]);
      }
      rc |= gemm_nt(dQKV, 3L * D, ly.Wqkv_o, 3 * D, D, dXin, D, units, s);
      if (rc) return jq_fail(JQ_EINVAL, "attention input adjoint launch rejected");
      // LN1 (pre-LN): dH = dH1 + LN1'(dXin); otherwise (U1 = H + ..., XIN = H) dH = dH1 + dXin
      if (pre) {
        launch_ln_bwd(units, D, Ht[l], D, ly.ln1_s, dXin, D, 1e-5f, dH1, D, gH, D, Pln, D, s);
        rc = wg(units, nullptr, 0, 1, Pln, D, D, ly.off_ln1s);
        rc |= wg(units, nullptr, 0, 1, dXin, D, D, ly.off_ln1b);
      } else {
        launch_add_rows(units, D, dH1, D, dXin, D, gH, D, s);
      }
    }
    if (rc) return jq_fail(JQ_EINVAL, "layer adjoint launch rejected");
    // input projection
    rc = wg(units, F0t, kin, 4 * M + 1, gH, D, D, off_win);
    rc |= wg(units, nullptr, 0, 1, gH, D, D, off_bin);
    if (d.jastrow) {
      launch_jastrow_grad(ec, Bc, N, n_up, jas[0], jas[1], cot, b0, jtmp, s);
      rc |= wg(Bc, nullptr, 0, 1, jtmp, 2, 2, off_jas);
    }
    if (rc) return jq_fail(JQ_EINVAL, "input adjoint launch rejected");
  }
  // the pack's q kernel and bias enter as 1/sqrt(dh) times the device's
  const double qs = 1.0 / std::sqrt((double)dh);
  for (const PsiLayer& ly : pl) {
    launch_scale_f64(gacc + ly.off_wq, (long)D * D, qs, s);
    launch_scale_f64(gacc + ly.off_bq, D, qs, s);
  }
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
