// Psiformer pipeline (app/molecule/wavefunction/psiformer.py:93-166,
// wavefunction/backbone/psiformer.py:29-187): weight packing and the
// per-chunk forward / forward-Laplacian sequence.
//
// Per layer, pre-LN (the default):  x += Wo . MHA(LN1(x));  x += MLP(LN2(x));
// post-LN:  x = LN1(x + Wo . MHA(x));  x = LN2(x + MLP(x));  null: no LayerNorm
// (backbone/psiformer.py:64-98); MLP(y) = tanh(W2 tanh(W1 y + b1) + b2).
// Attention per (walker w, head h), q pre-scaled by 1/sqrt(dh) in the weights:
//   S1[(i,c)][j] = sum_d Q[i][c][d] K[j][0][d]          (Q Jacobian/lap rows x key values)
//   S2[(j,c)][i] = sum_d K[j][c][d] Q[i][0][d]
//   S3[i][j]     = sum_{c>=2,d} Q[i][c][d] K[j][c][d]    (Laplacian cross term)
//   softmax_lap -> A1[(i,c)][j] (value, Laplacian, Jacobian of a), A2[i][(j,c)] = 2 J_a
//   O[(i,c)][d]  = sum_j A1[(i,c)][j] V[j][0][d] + [c >= 1] sum_j a[i][j] V[j][c][d]
//   O[(i,1)][d] += sum_{j,c>=2} 2 J_a[i][c][j] J_V[j][c][d]
// which is the product rule of laplacian/primitives/core.py:494-500 applied to
// q.k and a.v with dense Jacobians.

#include <algorithm>
#include <cmath>
#include <cstring>

#include "model.h"
#include "prof.h"


namespace {
bool all_zero(const float* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (p[i] != 0.f) return false;
  return true;
}

// dst (ncols x kpad) at column offset c0 <- transposed rows [r0, r0+nr) of W
// (rows x ldw_in, taking columns [cin0, cin0 + ncols)), scaled.
void put_t(std::vector<float>& dst, int kpad, const float* W, int ldw_in, int cin0, int ncols, int r0, int nr,
           int row_out0, float scale = 1.f) {
  for (int r = 0; r < nr; ++r)
    for (int n = 0; n < ncols; ++n)
      dst[(size_t)(row_out0 + n) * kpad + r] = scale * W[(size_t)(r0 + r) * ldw_in + cin0 + n];
}

constexpr size_t kBudget = 16ull << 30;
}  // namespace

int jq_model::init_psiformer(const float* packed, size_t n_floats) {
  L = d.n_layers;
  H = d.num_heads;
  dh = d.heads_dim;
  D = H * dh;
  if (L < 1 || H < 1 || dh < 4 || dh % 32 != 0 || d.n_mlp < 0 || d.n_mlp > 3 || (D != 64 && D != 128 && D != 256 && D != 512))
    return jq_fail(JQ_EINVAL, "unsupported Psiformer shape (heads_dim a multiple of 32, model dim 64/128/256/512)");
  if (Rp > 128) return jq_fail(JQ_EINVAL, "Psiformer supports N <= 42 electrons (tangent tile <= 128 rows)");
  PackReader rd{packed, n_floats};
  const int fin = 4 * M + 1;
  kin = round_up(fin, 32);
  const float* Wi = rd.take((size_t)fin * D);
  const float* bi = rd.take(D);
  if (!rd.ok) return jq_fail(JQ_EINVAL, "parameter pack too short (Psiformer input)");
  off_win = Wi - packed;
  off_bin = bi - packed;
  if ((options & JQ_OPT_NO_INPUT_BIAS) && !all_zero(bi, D))
    return jq_fail(JQ_EINVAL, "input_bias=False but the pack's input bias is not zero");
  {
    std::vector<float> wt((size_t)D * kin, 0.f);
    put_t(wt, kin, Wi, D, 0, D, 0, fin, 0);
    int rc = upload_wt(wt, D, &Win, &Win_lo);
    if (!rc) rc = upload(bi, 4L * D, (void**)&bin);
    if (rc) return rc;
  }
  const float qs = 1.f / std::sqrt((float)dh);
  pl.resize(L);
  for (int l = 0; l < L; ++l) {
    PsiLayer& ly = pl[l];
    const float* l1s = rd.take(D);
    const float* l1b = rd.take(D);
    const float* wq = rd.take((size_t)D * D);
    const float* bq = rd.take(D);
    const float* wk = rd.take((size_t)D * D);
    const float* bk = rd.take(D);
    const float* wv = rd.take((size_t)D * D);
    const float* bv = rd.take(D);
    const float* wo = rd.take((size_t)D * D);
    const float* bo = rd.take(D);
    const float* l2s = rd.take(D);
    const float* l2b = rd.take(D);
    ly.mdims.clear();
    ly.mdims.push_back(D);
    for (int j = 0; j < d.n_mlp; ++j) ly.mdims.push_back(d.mlp_hidden[j]);
    ly.mdims.push_back(D);
    std::vector<const float*> mw, mb;
    for (size_t j = 0; j + 1 < ly.mdims.size(); ++j) {
      mw.push_back(rd.take((size_t)ly.mdims[j] * ly.mdims[j + 1]));
      mb.push_back(rd.take(ly.mdims[j + 1]));
    }
    if (!rd.ok) return jq_fail(JQ_EINVAL, "parameter pack too short (Psiformer layer)");
    if ((options & JQ_OPT_NO_ATTN_BIAS) &&
        !(all_zero(bq, D) && all_zero(bk, D) && all_zero(bv, D) && all_zero(bo, D)))
      return jq_fail(JQ_EINVAL, "with_bias=False but the pack's attention biases are not zero");
    for (size_t j = 1; j + 1 < ly.mdims.size(); ++j)
      if (ly.mdims[j] < 1) return jq_fail(JQ_EINVAL, "bad MLP hidden width");
    int rc = upload(l1s, 4L * D, (void**)&ly.ln1_s);
    if (!rc) rc = upload(l1b, 4L * D, (void**)&ly.ln1_b);
    if (!rc) rc = upload(l2s, 4L * D, (void**)&ly.ln2_s);
    if (!rc) rc = upload(l2b, 4L * D, (void**)&ly.ln2_b);
    std::vector<float> wqkv((size_t)3 * D * D, 0.f), bqkv(3 * D);
    put_t(wqkv, D, wq, D, 0, D, 0, D, 0, qs);
    put_t(wqkv, D, wk, D, 0, D, 0, D, D);
    put_t(wqkv, D, wv, D, 0, D, 0, D, 2 * D);
    for (int n = 0; n < D; ++n) {
      bqkv[n] = qs * bq[n];
      bqkv[D + n] = bk[n];
      bqkv[2 * D + n] = bv[n];
    }
    if (!rc) rc = upload_wt(wqkv, 3 * D, &ly.Wqkv, &ly.Wqkv_lo);
    if (!rc) rc = upload(bqkv.data(), 12L * D, (void**)&ly.bqkv);
    std::vector<float> wot((size_t)D * D, 0.f);
    put_t(wot, D, wo, D, 0, D, 0, D, 0);
    if (!rc) rc = upload_wt(wot, D, &ly.Wo, &ly.Wo_lo);
    if (!rc) rc = upload(bo, 4L * D, (void**)&ly.bo);
    ly.Wm.resize(mw.size());
    ly.Wm_lo.resize(mw.size());
    ly.bm.resize(mw.size());
    for (size_t j = 0; j < mw.size() && !rc; ++j) {
      const int a = ly.mdims[j], b = ly.mdims[j + 1];
      const int ka = round_up(a, 32);
      std::vector<float> wt((size_t)b * ka, 0.f);
      put_t(wt, ka, mw[j], b, 0, b, 0, a, 0);
      rc = upload_wt(wt, b, &ly.Wm[j], &ly.Wm_lo[j]);
      if (!rc) rc = upload(mb[j], 4L * b, (void**)&ly.bm[j]);
    }
    if (rc) return rc;
    if (grads) {  // pack offsets and original-orientation FP32 weights
      ly.off_ln1s = l1s - packed;
      ly.off_ln1b = l1b - packed;
      ly.off_wq = wq - packed;
      ly.off_bq = bq - packed;
      ly.off_wk = wk - packed;
      ly.off_bk = bk - packed;
      ly.off_wv = wv - packed;
      ly.off_bv = bv - packed;
      ly.off_wo = wo - packed;
      ly.off_bo = bo - packed;
      ly.off_ln2s = l2s - packed;
      ly.off_ln2b = l2b - packed;
      std::vector<float> wo3((size_t)D * 3 * D);
      for (int r = 0; r < D; ++r)
        for (int n = 0; n < D; ++n) {
          wo3[(size_t)r * 3 * D + n] = qs * wq[(size_t)r * D + n];
          wo3[(size_t)r * 3 * D + D + n] = wk[(size_t)r * D + n];
          wo3[(size_t)r * 3 * D + 2 * D + n] = wv[(size_t)r * D + n];
        }
      rc = upload(wo3.data(), wo3.size() * 4, (void**)&ly.Wqkv_o);
      if (!rc) rc = upload(wo, 4L * D * D, (void**)&ly.Wo_o);
      ly.Wm_o.assign(mw.size(), nullptr);
      ly.off_mw.resize(mw.size());
      ly.off_mb.resize(mw.size());
      for (size_t j = 0; j < mw.size() && !rc; ++j) {
        const int a = ly.mdims[j], b = ly.mdims[j + 1], bp = round_up(b, 32);
        std::vector<float> w2((size_t)a * bp, 0.f);
        for (int r = 0; r < a; ++r)
          for (int n = 0; n < b; ++n) w2[(size_t)r * bp + n] = mw[j][(size_t)r * b + n];
        rc = upload(w2.data(), w2.size() * 4, (void**)&ly.Wm_o[j]);
        ly.off_mw[j] = mw[j] - packed;
        ly.off_mb[j] = mb[j] - packed;
      }
      if (rc) return rc;
    }
  }
  // outputs (orbital projection, envelope), then the Jastrow
  const size_t pos_out = rd.pos;
  {
    int rc = read_outputs(rd, D);
    if (rc) return rc;
  }
  const float* jas_p = rd.take(2);
  if (!jas_p) return jq_fail(JQ_EINVAL, "parameter pack too short (Jastrow)");
  if (grads) {  // output offsets (orbital kernels, orbital biases, envelopes) and the Jastrow
    const long KN = (long)K * N;
    ldko = round_up((int)KN, 32);
    const long ne = (long)N * M * K;
    const long nb = (options & JQ_OPT_ORBITAL_BIAS) ? (long)norb * KN : 0;
    for (int c = 0; c < norb; ++c) {
      off_orb[c] = (long)pos_out + c * D * KN;
      off_borb[c] = (long)pos_out + norb * D * KN + c * KN;
      off_env[c] = (long)pos_out + norb * D * KN + nb + c * 2 * ne;
      std::vector<float> wo((size_t)D * ldko, 0.f);
      for (int r = 0; r < D; ++r)
        for (long n = 0; n < KN; ++n) wo[(size_t)r * ldko + n] = packed[off_orb[c] + r * KN + n];
      int rc = upload(wo.data(), wo.size() * 4, (void**)&Worb_o[c]);
      if (rc) return rc;
    }
    if (norb == 1) Worb_o[1] = Worb_o[0];
    off_jas = jas_p - packed;
  }
  jas[0] = jas_p[0];
  jas[1] = jas_p[1];
  if (rd.pos != n_floats) return jq_fail(JQ_EINVAL, "parameter pack size mismatch");
  n_params = (long)n_floats;

  // ---- workspace ----
  int hmax = D;
  for (auto& ly : pl)
    for (int v : ly.mdims) hmax = std::max(hmax, round_up(v, 32));
  const long NpS = round_up(N, 4), Np32 = round_up(N, 32);
  // Laplacian-mode attention runs fused (no score tensors in HBM) when supported
  const bool fused_attn = attn_lap_supported(N, dh);
  auto attn_scratch = [&](long R) -> long {  // floats per walker for the score tensors
    if (R > 1 && fused_attn) return 0;
    const long u = (long)N * R;
    const long lda2 = (long)round_up((int)(N * R), 32);
    return 2L * H * u * NpS + (long)H * N * NpS + (long)H * u * Np32 + (R > 1 ? (long)H * N * lda2 : 0);
  };
  auto per_walker = [&](long R) {
    const long u = (long)N * R;
    return (size_t)4 * (u * D + u * hmax + u * std::max<long>(hmax, kin) + 3 * u * D + u * (D + 4) +
                        attn_scratch(R) + u * round_up(K * N, 4) + (long)N * K * N * 5) +
           (size_t)8 * K * (3 + T) + 64;
  };
  ldphi = round_up(K * N, 4);
  const size_t pw_lap = per_walker(Rp), pw_fwd = per_walker(1);
  chunk = d.chunk_walkers > 0 ? d.chunk_walkers : (int)std::max<size_t>(1, kBudget / pw_lap);
  chunk = std::min(chunk, 1 << 16);
  chunk_fwd = (int)std::min<size_t>((size_t)chunk * pw_lap / pw_fwd, 1 << 18);
  auto nmax = [&](long per_lap, long per_fwd) {
    return (size_t)std::max((long)chunk * per_lap, (long)chunk_fwd * per_fwd);
  };
  const long uL = (long)N * Rp, uF = N;
  P[0] = alloc(4 * nmax(uL * hmax, uF * hmax));                                 // x (swaps with T1 post-LN)
  P[1] = alloc(4 * nmax(uL * hmax, uF * hmax));                                 // T1
  P[2] = alloc(4 * nmax(uL * std::max(hmax, kin), uF * std::max(hmax, kin)));   // T2 / X0
  P[3] = alloc(4 * nmax(3 * uL * D, 3 * uF * D));                               // QKV (head-split)
  P[4] = alloc(4 * nmax(uL * (D + 4), uF * (D + 4)));                           // attention output (padded rows)
  P[5] = alloc(4 * std::max<size_t>(1, nmax(attn_scratch(Rp), attn_scratch(1))));  // scores / softmax
  Phi = (float*)alloc(4 * nmax(uL * ldphi, uF * ldphi));
  env = (float*)alloc(4 * (size_t)std::max(chunk, chunk_fwd) * N * K * N * 5);
  det = (double*)alloc(8 * nmax((long)K * (3 + T), (long)K * 3));
  for (int i = 0; i < 6; ++i)
    if (!P[i]) return jq_fail(JQ_ENOMEM, "workspace allocation failed");
  if (!Phi || !env || !det) return jq_fail(JQ_ENOMEM, "workspace allocation failed");
  int rc = alloc_mh_scratch();
  if (!rc && grads) rc = alloc_grad_workspace_psi(4ull << 30);
  return rc;
}

int jq_model::forward_psiformer(const float* elec, int Bc, bool lap, hipStream_t s) {
  const FermiDims fd = dims(Bc, lap);
  const int RP = fd.Rp;
  const int R = lap ? T + 2 : 1;
  const long units = (long)Bc * N;
  float* x = (float*)P[0];
  float* T1 = (float*)P[1];
  float* T2 = (float*)P[2];
  float* QKV = (float*)P[3];
  float* Om = (float*)P[4];
  // Attention output rows are D + 4 floats apart: its Laplacian/Jacobian
  // blocks are written as 256-byte segments one unit (Rp rows) apart, and a
  // power-of-two unit stride would send them all to the same HBM channels.
  const int DO = D + 4;
  const int NpS = round_up(N, 4), Np32 = round_up(N, 32);
  const long lda2 = round_up(N * RP, 32);
  float* S1 = (float*)P[5];
  float* S2 = S1 + (long)Bc * H * N * RP * NpS;
  float* S3 = S2 + (long)Bc * H * N * RP * NpS;
  float* A1 = S3 + (long)Bc * H * N * NpS;
  float* A2 = A1 + (long)Bc * H * N * RP * Np32;

  // Pre-LN layers on the Laplacian path: where a GEMM's workgroup holds whole
  // units and every feature, the LayerNorm that follows the residual update
  // runs in that GEMM's epilogue (Yln) instead of as a kernel of its own.
  const bool lnf = lap && ln_mode == JQ_LN_PRE && gemm_ln_fusable(RP, D, prec);
  auto gemm = [&](const float* X, long ldx, void* Wt, void* Wlo, long wplane, int K_, float* Y, long ldy, int ncols,
                  const float* bias, int act, const float* Res, int zpad = 0, float* yln = nullptr,
                  const float* lns = nullptr, const float* lnb = nullptr) {
    GemmArgs g{};
    g.X = X;
    g.ldx = ldx;
    g.Wt = Wt;
    g.Wt_lo = Wlo;
    g.wplane = wplane;
    g.ldw = K_;
    g.Y = Y;
    g.ldy = ldy;
    g.rows = (int)(units * RP);
    g.K = K_;
    g.ncols = ncols;
    g.bias = bias;
    g.R = RP;
    g.act = act;
    g.Res = Res;
    g.ldres = ldy;
    g.res_scale = 1.f;
    g.Bc = Bc;
    g.n_up = n_up;
    g.n_dn = n_dn;
    g.wmaj = 1;
    g.zpad = zpad;
    g.Yln = yln;
    g.ln_s = lns;
    g.ln_b = lnb;
    g.ln_eps = 1e-5f;
    g.ln_R = R;
    ProfScope ps(lap ? PROF_GEMM_LAYER_LAP : -1, s);
    return launch_gemm(g, prec, false, s);
  };

  int rc = launch_psi_features(fd, elec, atoms, T2, kin, (options & JQ_OPT_NO_RESCALE) ? 0 : 1, s);
  // ln1_in_t1: T1 already holds LN1(x) of the coming layer (fused above)
  bool ln1_in_t1 = lnf && L > 0;
  rc |= gemm(T2, kin, Win, Win_lo, (long)D * kin, kin, x, D, D, bin, ACT_NONE, nullptr, 0, ln1_in_t1 ? T1 : nullptr,
             L > 0 ? pl[0].ln1_s : nullptr, L > 0 ? pl[0].ln1_b : nullptr);
  for (int l = 0; l < L && !rc; ++l) {
    const PsiLayer& ly = pl[l];
    const float* qin = x;  // attention input: LN1(x) pre-LN, x otherwise
    if (ln_mode == JQ_LN_PRE) {
      if (!ln1_in_t1) rc |= launch_layernorm_lap((int)units, RP, R, D, x, T1, ly.ln1_s, ly.ln1_b, 1e-5f, s);
      qin = T1;
    }
    ln1_in_t1 = false;
    {
      GemmArgs g{};
      g.X = qin;
      g.ldx = D;
      g.Wt = ly.Wqkv;
      g.Wt_lo = ly.Wqkv_lo;
      g.wplane = 3L * D * D;
      g.ldw = D;
      g.Y = QKV;
      g.ldy = 3 * D;
      g.rows = (int)(units * RP);
      g.K = D;
      g.ncols = 3 * D;
      g.bias = ly.bqkv;
      g.R = RP;
      g.act = ACT_NONE;
      g.res_scale = 1.f;
      g.wmaj = 1;
      g.hs_H = H;
      g.hs_dh = dh;
      g.hs_D = D;
      g.hs_N = N;
      ProfScope ps(lap ? PROF_GEMM_LAYER_LAP : -1, s);
      rc |= launch_gemm(g, prec, false, s);
    }
    // ---- attention ----
    if (lap && attn_lap_supported(N, dh)) {
      ProfScope ps(PROF_ATTN_LAP, s);
      AttnArgs aa{};
      aa.QKV = QKV;
      aa.O = Om;
      aa.ldo = DO;
      aa.N = N;
      aa.Rp = RP;
      aa.R = R;
      aa.H = H;
      aa.dh = dh;
      aa.nz = Bc * H;
      aa.h3 = prec == PREC_FP16X3 ? 1 : 0;
      rc |= launch_attn_lap(aa, s);
    } else if (!lap && attn_fwd_supported(N, dh)) {
      AttnArgs aa{};
      aa.QKV = QKV;
      aa.O = Om;
      aa.ldo = DO;
      aa.N = N;
      aa.Rp = 1;
      aa.R = 1;
      aa.H = H;
      aa.dh = dh;
      aa.nz = Bc * H;
      rc |= launch_attn_fwd(aa, s);
    } else {
      ProfScope ps(lap ? PROF_ATTN_LAP : -1, s);
      const long blk = (long)N * RP * dh;   // one (w, part, h) block of QKV
      const long sw = 3L * H * blk;
      const float* Q = QKV;
      const float* Kx = QKV + (long)H * blk;
      const float* V = QKV + 2L * H * blk;
      BGemmArgs b{};
      // S1 = Q_rows . K_values
      b = BGemmArgs{};
      b.X = Q; b.ldx = dh; b.sxw = sw; b.sxh = blk;
      b.W = Kx; b.ldw = (long)RP * dh; b.sww = sw; b.swh = blk;
      b.Y = S1; b.ldy = NpS; b.syw = (long)H * N * RP * NpS; b.syh = (long)N * RP * NpS;
      b.rows = N * RP; b.K = dh; b.ncols = N; b.kvalid = dh; b.nw = Bc; b.nh = H; b.nn = 0;
      rc |= launch_bgemm(b, prec, s);
      if (lap) {
        // S2 = K_rows . Q_values
        b.X = Kx;
        b.W = Q;
        b.Y = S2;
        rc |= launch_bgemm(b, prec, s);
        // S3 = cross term over Jacobian rows
        BGemmArgs c3{};
        c3.X = Q + 2 * dh; c3.ldx = (long)RP * dh; c3.sxw = sw; c3.sxh = blk;
        c3.W = Kx + 2 * dh; c3.ldw = (long)RP * dh; c3.sww = sw; c3.swh = blk;
        c3.Y = S3; c3.ldy = NpS; c3.syw = (long)H * N * NpS; c3.syh = (long)N * NpS;
        c3.rows = N; c3.K = (RP - 2) * dh; c3.ncols = N; c3.kvalid = 0; c3.nw = Bc; c3.nh = H; c3.nn = 0;
        rc |= launch_bgemm(c3, prec, s);
      }
      SoftmaxArgs sa{};
      sa.S1 = S1;
      sa.S2 = S2;
      sa.S3 = S3;
      sa.lds = NpS;
      sa.A1 = A1;
      sa.lda1 = Np32;
      sa.A2 = lap ? A2 : nullptr;
      sa.N = N;
      sa.Rp = RP;
      sa.R = R;
      sa.nz = Bc * H;
      rc |= launch_softmax_lap(sa, s);
      // O = A1 . V_values   (all rows of a; NN)
      BGemmArgs g4{};
      g4.X = A1; g4.ldx = Np32; g4.sxw = (long)H * N * RP * Np32; g4.sxh = (long)N * RP * Np32;
      g4.W = V; g4.ldw = (long)RP * dh; g4.sww = sw; g4.swh = blk;
      g4.Y = Om; g4.ldy = DO; g4.syw = (long)N * RP * DO; g4.syh = dh;
      g4.rows = N * RP; g4.K = Np32; g4.ncols = dh; g4.kvalid = N; g4.nw = Bc; g4.nh = H; g4.nn = 1;
      rc |= launch_bgemm(g4, prec, s);
      if (lap) {
        // O[i][(c,d)] += a[i][:] . V[:, (c,d)] for the Laplacian and Jacobian rows c >= 1
        // (the value row a.v came from the previous product)
        BGemmArgs g5{};
        g5.X = A1; g5.ldx = (long)RP * Np32; g5.sxw = g4.sxw; g5.sxh = g4.sxh;
        g5.W = V + dh; g5.ldw = (long)RP * dh; g5.sww = sw; g5.swh = blk;
        g5.Y = Om + DO; g5.ldy = (long)RP * DO; g5.syw = g4.syw; g5.syh = dh; g5.Res = Om + DO;
        g5.rows = N; g5.K = Np32; g5.ncols = (RP - 1) * dh; g5.kvalid = N; g5.nw = Bc; g5.nh = H; g5.nn = 1;
        g5.ycw = dh; g5.ycs = DO;
        rc |= launch_bgemm(g5, prec, s);
        // O[i][1][d] += sum_{j,c} 2 J_a[i][(j,c)] V[(j,c)][d]
        BGemmArgs g6{};
        g6.X = A2; g6.ldx = lda2; g6.sxw = (long)H * N * lda2; g6.sxh = (long)N * lda2;
        g6.W = V; g6.ldw = dh; g6.sww = sw; g6.swh = blk;
        g6.Y = Om + DO; g6.ldy = (long)RP * DO; g6.syw = g4.syw; g6.syh = dh; g6.Res = Om + DO;
        g6.rows = N; g6.K = (int)lda2; g6.ncols = dh; g6.kvalid = N * RP; g6.nw = Bc; g6.nh = H; g6.nn = 1;
        rc |= launch_bgemm(g6, prec, s);
      }
    }
    // x += Wo O + bo  (and T1 = LN2(x) when fused)
    rc |= gemm(Om, DO, ly.Wo, ly.Wo_lo, (long)D * D, D, x, D, D, ly.bo, ACT_NONE, x, 0, lnf ? T1 : nullptr,
               ly.ln2_s, ly.ln2_b);
    if (ln_mode == JQ_LN_POST) {  // x = LN1(x + attention)
      rc |= launch_layernorm_lap((int)units, RP, R, D, x, T1, ly.ln1_s, ly.ln1_b, 1e-5f, s);
      std::swap(x, T1);
    }
    // MLP on LN2(x) (pre-LN) or x, residual into x
    float* cur = x;
    if (ln_mode == JQ_LN_PRE) {
      if (!lnf) rc |= launch_layernorm_lap((int)units, RP, R, D, x, T1, ly.ln2_s, ly.ln2_b, 1e-5f, s);
      cur = T1;
    }
    int curw = D;
    const int nm = (int)ly.Wm.size();
    for (int j = 0; j < nm && !rc; ++j) {
      const int outw = ly.mdims[j + 1];
      const int ka = round_up(curw, 32);
      const bool last = j == nm - 1;
      float* dst = last ? x : (cur == T1 ? T2 : T1);
      // the last one writes LN1(x) of the next layer into T1 when fused.  If
      // its input is T1 itself (one MLP layer), the layouts agree and every
      // tile holds whole units and all columns, so each tile overwrites only
      // rows it has finished reading.
      const bool fuse_next = lnf && last && l + 1 < L && (cur != T1 || ka == D);
      // hidden outputs: padding columns up to the next K multiple are zeroed
      // for the following GEMM
      rc |= gemm(cur, ka, ly.Wm[j], ly.Wm_lo[j], (long)outw * ka, ka, dst, last ? D : round_up(outw, 32), outw,
                 ly.bm[j], ACT_TANH, last ? x : nullptr, last ? 0 : round_up(outw, 32), fuse_next ? T1 : nullptr,
                 fuse_next ? pl[l + 1].ln1_s : nullptr, fuse_next ? pl[l + 1].ln1_b : nullptr);
      ln1_in_t1 = fuse_next;
      cur = dst;
      curw = outw;
    }
    if (ln_mode == JQ_LN_POST) {  // x = LN2(x + MLP)
      rc |= launch_layernorm_lap((int)units, RP, R, D, x, T1, ly.ln2_s, ly.ln2_b, 1e-5f, s);
      std::swap(x, T1);
    }
  }
  if (rc) return jq_fail(JQ_EINVAL, "Psiformer launch rejected");
  return psi_outputs(elec, x, Bc, lap, s);
}

int jq_model::psi_outputs(const float* elec, const float* x, int Bc, bool lap, hipStream_t s) {
  const FermiDims fd = dims(Bc, lap);
  const int RP = fd.Rp;
  // orbitals: per spin channel, batched over walkers (walker-major units)
  for (int c = 0; c < nch; ++c) {
    const int i0 = (nch == 1 || c == 0) ? 0 : n_up;
    const int ns = nch == 1 ? N : (c == 0 ? n_up : n_dn);
    GemmArgs go{};
    go.X = x + (long)i0 * RP * D;
    go.ldx = D;
    go.Wt = Worb[c];
    go.Wt_lo = Worb_lo[c];
    go.wplane = (long)K * N * korb;
    go.ldw = korb;
    go.Y = Phi + (long)i0 * RP * ldphi;
    go.ldy = ldphi;
    go.rows = ns * RP;
    go.K = korb;
    go.ncols = K * N;
    go.R = 1;
    go.lrp = lap ? RP : 0;
    go.act = ACT_NONE;
    go.bias = borb[c];
    if (lap) {  // unit-tiled launch (the bias, if any, on the value rows only)
      go.R = RP;
      go.wmaj = 1;
      go.Bc = Bc;
      go.n_up = n_up;
      go.n_dn = n_dn;
    }
    if (lap) {  // a walker's channel rows are whole 128-row tiles: batch over walkers
      go.nbatch = Bc;
      go.sxb = (long)N * RP * D;
      go.syb = (long)N * RP * ldphi;
    } else {  // value rows: one launch over the walkers' channel rows as row groups
      go.rows = Bc * ns;
      go.rgrp = ns;
      go.sxg = (long)N * D;
      go.syg = (long)N * ldphi;
    }
    ProfScope ps(lap ? PROF_GEMM_ORB_LAP : -1, s);
    if (launch_gemm(go, prec, false, s)) return jq_fail(JQ_EINVAL, "orbital GEMM launch rejected");
  }
  const float* pis[2] = {pi[0], pi[1]};
  const float* sgs[2] = {sigma[0], sigma[1]};
  launch_envelope(fd, elec, atoms, pis, sgs, d.envelope, env, s);
  if (skip_det) return 0;
  ProfScope ps(lap ? PROF_DET_LAP : -1, s);
  if (launch_det(fd, Phi, ldphi, env, det, s)) return jq_fail(JQ_EINVAL, "determinant launch rejected");
  return 0;
}
