// Model construction (parameter pack -> device layouts) and the per-chunk
// pipelines shared by log|psi| and the local energy.

#include <algorithm>
#include <cmath>
#include <cstring>

#include "model.h"
#include "prof.h"

int gemm_unit_rows(int R);

namespace {

// Transpose rows [r0, r0 + nr) of a row-major (rows x ncols) matrix into
// dst (ncols x kpad) at column offset c0.
void put_t(std::vector<float>& dst, int kpad, const float* W, int ncols, int r0, int nr, int c0) {
  for (int r = 0; r < nr; ++r)
    for (int n = 0; n < ncols; ++n) dst[(size_t)n * kpad + c0 + r] = W[(size_t)(r0 + r) * ncols + n];
}

constexpr size_t kDefaultBudget = 16ull << 30;  // bytes of walker-chunk workspace

}  // namespace

int jq_model::init(const jq_model_desc& desc, const float* packed, size_t n_floats) {
  d = desc;
  kind = desc.kind;
  n_up = desc.n_up;
  n_dn = desc.n_dn;
  N = n_up + n_dn;
  T = 3 * N;
  M = desc.n_atoms;
  K = desc.ndets;
  prec = desc.precision;
  xbf = false;
  nch = (n_up > 0 && n_dn > 0) ? 2 : 1;
  if (N < 1 || N > 63 || M < 1 || K < 1 || n_up < 0 || n_dn < 0)
    return jq_fail(JQ_EINVAL, "unsupported system size (1 <= N <= 63 electrons, >= 1 atom, >= 1 det)");
  if (prec < 0 || prec > JQ_PREC_FP16X3) return jq_fail(JQ_EINVAL, "bad precision");
  if (desc.envelope == JQ_ENV_DIAGONAL)
    return jq_fail(JQ_EINVAL, "the diagonal envelope (EnvelopeType.diagonal) is not supported");
  if (desc.envelope < 0 || desc.envelope > 2) return jq_fail(JQ_EINVAL, "bad envelope kind");
  if (!desc.atoms || !desc.charges) return jq_fail(JQ_EINVAL, "atoms / charges missing");
  // module options (checked before any allocation)
  options = desc.options;
  ln_mode = desc.layer_norm_mode;
  const int known = JQ_OPT_NO_ATTN_BIAS | JQ_OPT_NO_INPUT_BIAS | JQ_OPT_ORBITAL_BIAS | JQ_OPT_NO_RESCALE |
                    JQ_OPT_SHARED_ORBITALS | JQ_OPT_USE_LAST_LAYER | JQ_OPT_WITH_GRADIENTS;
  if (options & ~known) return jq_fail(JQ_EINVAL, "unknown option bits");
  if (options & JQ_OPT_USE_LAST_LAYER)
    return jq_fail(JQ_EINVAL, "FermiNet use_last_layer=True (aggregated orbital input) is not supported");
  if (kind == JQ_KIND_FERMINET) {
    if (ln_mode != JQ_LN_PRE) return jq_fail(JQ_EINVAL, "layer_norm_mode applies to the Psiformer only");
    if (options & (JQ_OPT_NO_ATTN_BIAS | JQ_OPT_NO_INPUT_BIAS | JQ_OPT_ORBITAL_BIAS))
      return jq_fail(JQ_EINVAL, "attention / input / orbital bias options apply to the Psiformer only");
  } else if (ln_mode < JQ_LN_PRE || ln_mode > JQ_LN_NULL) {
    return jq_fail(JQ_EINVAL, "bad layer_norm_mode");
  }
  grads = (options & JQ_OPT_WITH_GRADIENTS) != 0;
  if (grads && kind == JQ_KIND_PSIFORMER) {
    if (!attn_fwd_supported(desc.n_up + desc.n_dn, desc.heads_dim))
      return jq_fail(JQ_EINVAL, "Psiformer parameter gradients need N <= 64 and heads_dim 32 or 64");
  }
  if (grads && kind == JQ_KIND_FERMINET)
    for (int l = 0; l < desc.n_layers && l < JQ_MAX_LAYERS; ++l)
      if (desc.hidden_single[l] % 32 != 0)
        return jq_fail(JQ_EINVAL, "parameter gradients need one-electron widths that are multiples of 32");
  norb = (nch == 2 && !(options & JQ_OPT_SHARED_ORBITALS)) ? 2 : 1;
  Rp = gemm_unit_rows(T + 2);
  if (Rp < 0) return jq_fail(JQ_EINVAL, "too many electrons for the tangent tile");
  atoms_h.assign(desc.atoms, desc.atoms + 3 * M);
  charges_h.assign(desc.charges, desc.charges + M);
  int rc = upload(atoms_h.data(), 12 * M, (void**)&atoms);
  if (rc) return rc;
  rc = upload(charges_h.data(), 4 * M, (void**)&charges);
  if (rc) return rc;
  if (kind == JQ_KIND_FERMINET) return init_ferminet(packed, n_floats);
  if (kind == JQ_KIND_PSIFORMER) return init_psiformer(packed, n_floats);
  return jq_fail(JQ_EINVAL, "unknown model kind");
}

// Output layers shared by both ansatzes: the orbital projection, spin-split
// (SplitChannelDense) or shared (one DenseGeneral), with an optional bias
// (orbital.py:10-78), and the envelope parameters per orbital kernel
// (envelope.py:58-96).  Pack order: the norb kernels (d_last, K*N), then the
// norb biases (K*N) when bias_orbitals, then (pi, sigma) per envelope.
int jq_model::read_outputs(PackReader& rd, int d_last) {
  jq_model* m = this;
  m->korb = round_up(d_last, 32);
  const int ncol = m->K * m->N;
  for (int c = 0; c < m->norb; ++c) {
    const float* W = rd.take((size_t)d_last * ncol);
    if (!W) return jq_fail(JQ_EINVAL, "parameter pack too short (orbital)");
    std::vector<float> wt((size_t)ncol * m->korb, 0.f);
    put_t(wt, m->korb, W, ncol, 0, d_last, 0);
    int rc = m->upload_wt(wt, ncol, &m->Worb[c], &m->Worb_lo[c]);
    if (rc) return rc;
  }
  if (m->options & JQ_OPT_ORBITAL_BIAS) {
    for (int c = 0; c < m->norb; ++c) {
      const float* b = rd.take(ncol);
      if (!b) return jq_fail(JQ_EINVAL, "parameter pack too short (orbital bias)");
      int rc = m->upload(b, 4L * ncol, (void**)&m->borb[c]);
      if (rc) return rc;
    }
  }
  const size_t ne = (size_t)m->N * m->M * m->K;
  for (int c = 0; c < m->norb; ++c) {
    const float* pi = rd.take(ne);
    const float* sg = rd.take(ne);
    if (!pi || !sg) return jq_fail(JQ_EINVAL, "parameter pack too short (envelope)");
    const std::vector<float> pt = envelope_device_layout(pi, m->N, m->M, m->K);
    const std::vector<float> st = envelope_device_layout(sg, m->N, m->M, m->K);
    int rc = m->upload(pt.data(), ne * 4, (void**)&m->pi[c]);
    if (rc) return rc;
    rc = m->upload(st.data(), ne * 4, (void**)&m->sigma[c]);
    if (rc) return rc;
  }
  if (m->norb == 1) {  // one shared set serves both spin channels
    m->Worb[1] = m->Worb[0];
    m->Worb_lo[1] = m->Worb_lo[0];
    m->borb[1] = m->borb[0];
    m->pi[1] = m->pi[0];
    m->sigma[1] = m->sigma[0];
  }
  return 0;
}

int jq_model::init_ferminet(const float* packed, size_t n_floats) {
  L = d.n_layers;
  if (L < 1 || L > JQ_MAX_LAYERS) return jq_fail(JQ_EINVAL, "bad n_layers");
  PackReader rd{packed, n_floats};
  fl.resize(L);
  int f1 = 4 * M, f2 = 4;
  std::vector<const float*> Ws(L), W2s(L), bs(L), b2s(L);
  for (int l = 0; l < L; ++l) {
    FermiLayer& ly = fl[l];
    ly.f1 = f1;
    ly.f2 = f2;
    ly.d1 = d.hidden_single[l];
    ly.d2 = d.hidden_double[l];
    ly.has_pair = l < L - 1;
    ly.res1 = f1 == ly.d1;
    ly.res2 = ly.has_pair && f2 == ly.d2;
    ly.kmain = round_up(f1 + nch * f2, 32);
    ly.kg = round_up(nch * f1, 32);
    if (ly.d1 < 1 || ly.d2 < 1) return jq_fail(JQ_EINVAL, "bad hidden width");
    const int in = f1 * (1 + nch) + f2 * nch;
    ly.off_w = (long)rd.pos;
    Ws[l] = rd.take((size_t)in * ly.d1);
    ly.off_b = (long)rd.pos;
    bs[l] = rd.take(ly.d1);
    f1 = ly.d1;
    if (ly.has_pair) f2 = ly.d2;
  }
  for (int l = 0; l < L - 1; ++l) {
    fl[l].off_w2 = (long)rd.pos;
    W2s[l] = rd.take((size_t)fl[l].f2 * fl[l].d2);
    fl[l].off_b2 = (long)rd.pos;
    b2s[l] = rd.take(fl[l].d2);
  }
  if (!rd.ok) return jq_fail(JQ_EINVAL, "parameter pack too short (FermiNet layers)");
  for (int l = 0; l < L; ++l) {
    FermiLayer& ly = fl[l];
    // W rows: [h (f1) | g_one (nch*f1) | g_two (nch*f2)]  (aggregate_features order)
    std::vector<float> wm((size_t)ly.d1 * ly.kmain, 0.f), wg((size_t)ly.d1 * ly.kg, 0.f);
    put_t(wm, ly.kmain, Ws[l], ly.d1, 0, ly.f1, 0);
    put_t(wm, ly.kmain, Ws[l], ly.d1, ly.f1 * (1 + nch), nch * ly.f2, ly.f1);
    put_t(wg, ly.kg, Ws[l], ly.d1, ly.f1, nch * ly.f1, 0);
    int rc = upload_wt(wm, ly.d1, &ly.Wm, &ly.Wm_lo);
    if (!rc) rc = upload_wt(wg, ly.d1, &ly.Wg, &ly.Wg_lo);
    if (!rc) rc = upload(bs[l], 4 * ly.d1, (void**)&ly.b);
    if (!rc && grads) {  // original orientation for dX = dz W^T: rows [h | pair means], rows [spin means]
      std::vector<float> wmo((size_t)ly.kmain * ly.d1, 0.f), wgo((size_t)ly.kg * ly.d1, 0.f);
      for (int r = 0; r < ly.f1; ++r)
        for (int n = 0; n < ly.d1; ++n) wmo[(size_t)r * ly.d1 + n] = Ws[l][(size_t)r * ly.d1 + n];
      for (int r = 0; r < nch * ly.f2; ++r)
        for (int n = 0; n < ly.d1; ++n)
          wmo[(size_t)(ly.f1 + r) * ly.d1 + n] = Ws[l][(size_t)(ly.f1 * (1 + nch) + r) * ly.d1 + n];
      for (int r = 0; r < nch * ly.f1; ++r)
        for (int n = 0; n < ly.d1; ++n) wgo[(size_t)r * ly.d1 + n] = Ws[l][(size_t)(ly.f1 + r) * ly.d1 + n];
      rc = upload(wmo.data(), wmo.size() * 4, (void**)&ly.Wm_o);
      if (!rc) rc = upload(wgo.data(), wgo.size() * 4, (void**)&ly.Wg_o);
    }
    if (!rc && ly.has_pair) {
      rc = upload(W2s[l], 4L * ly.f2 * ly.d2, (void**)&ly.W2);
      if (!rc) rc = upload(b2s[l], 4L * ly.d2, (void**)&ly.b2);
    }
    if (rc) return rc;
  }
  const int d_last = fl[L - 1].d1;
  if (d_last % 32 != 0) return jq_fail(JQ_EINVAL, "last one-electron width must be a multiple of 32");
  const size_t pos_out = rd.pos;
  int rc = read_outputs(rd, d_last);
  if (rc) return rc;
  if (rd.pos != n_floats) return jq_fail(JQ_EINVAL, "parameter pack size mismatch");
  n_params = (long)n_floats;
  if (grads) {  // pack offsets of the outputs and the original-orientation orbital kernels
    const long KN = (long)K * N;
    ldko = round_up((int)KN, 32);
    const long ne = (long)N * M * K;
    for (int c = 0; c < norb; ++c) {
      off_orb[c] = (long)pos_out + c * d_last * KN;
      off_env[c] = (long)pos_out + norb * d_last * KN + c * 2 * ne;
      std::vector<float> wo((size_t)d_last * ldko, 0.f);
      for (int r = 0; r < d_last; ++r)
        for (long n = 0; n < KN; ++n) wo[(size_t)r * ldko + n] = packed[off_orb[c] + r * KN + n];
      rc = upload(wo.data(), wo.size() * 4, (void**)&Worb_o[c]);
      if (rc) return rc;
    }
    if (norb == 1) Worb_o[1] = Worb_o[0];
  }

  // ---- workspace ----
  size_t kx = 0, kgm = 0, d1m = 0, f2m = 4;
  for (auto& ly : fl) {
    kx = std::max(kx, (size_t)ly.kmain);
    kgm = std::max(kgm, (size_t)ly.kg);
    d1m = std::max(d1m, (size_t)ly.d1);
    f2m = std::max(f2m, (size_t)ly.d2);
  }
  kx = std::max(kx, (size_t)korb);
  ldphi = round_up(K * N, 4);
  auto per_walker = [&](int R, int RP2) {
    return 2 * (size_t)N * R * kx * es() + (size_t)R * kgm * es() + (size_t)R * d1m * 4 +
           2 * (size_t)N * N * RP2 * f2m * 4 + (size_t)N * R * ldphi * 4 + (size_t)N * K * N * 5 * 4 +
           (size_t)K * (3 + T) * 8 + 64;
  };
  const size_t pw_lap = per_walker(Rp, 8), pw_fwd = per_walker(1, 1);
  chunk = d.chunk_walkers > 0 ? d.chunk_walkers : (int)std::max<size_t>(1, kDefaultBudget / pw_lap);
  chunk = std::min(chunk, 1 << 16);
  chunk_fwd = (int)std::min<size_t>((size_t)chunk * pw_lap / pw_fwd, 1 << 20);
  const size_t nX = std::max((size_t)chunk * N * Rp * kx, (size_t)chunk_fwd * N * kx);
  const size_t ng = std::max((size_t)chunk * Rp * kgm, (size_t)chunk_fwd * kgm);
  const size_t nG = std::max((size_t)chunk * Rp * d1m, (size_t)chunk_fwd * d1m);
  const size_t nH = std::max((size_t)chunk * N * N * 8 * f2m, (size_t)chunk_fwd * N * N * f2m);
  const size_t nP = std::max((size_t)chunk * N * Rp * ldphi, (size_t)chunk_fwd * N * ldphi);
  const size_t nE = (size_t)std::max(chunk, chunk_fwd) * N * K * N * 5;
  const size_t nD = std::max((size_t)chunk * K * (3 + T), (size_t)chunk_fwd * K * 3);
  X[0] = alloc(nX * es());
  X[1] = alloc(nX * es());
  g = alloc(ng * es());
  G = (float*)alloc(nG * 4);
  H2[0] = (float*)alloc(nH * 4);
  H2[1] = (float*)alloc(nH * 4);
  Phi = (float*)alloc(nP * 4);
  env = (float*)alloc(nE * 4);
  det = (double*)alloc(nD * 8);
  if (!X[0] || !X[1] || !g || !G || !H2[0] || !H2[1] || !Phi || !env || !det)
    return jq_fail(JQ_ENOMEM, "workspace allocation failed");
  rc = alloc_mh_scratch();
  if (!rc && grads) rc = alloc_grad_workspace(4ull << 30);
  return rc;
}

// Proposals of one forward chunk and log|psi| of one chunk (either mode).
int jq_model::alloc_mh_scratch() {
  prop = (float*)alloc((size_t)chunk_fwd * T * 4);
  lp = (float*)alloc((size_t)std::max(chunk, chunk_fwd) * 4);
  if (!prop || !lp) return jq_fail(JQ_ENOMEM, "Metropolis scratch allocation failed");
  return 0;
}

// ---- FermiNet pipeline --------------------------------------------------------
int jq_model::forward_ferminet(const float* elec, int Bc, bool lap, hipStream_t s) {
  const FermiDims fd = dims(Bc, lap);
  const int RP = fd.Rp;
  void* Xc = X[0];
  void* Xn = X[1];
  int rc = launch_ferminet_features(fd, elec, atoms, Xc, fl[0].kmain, xbf, H2[0], s);
  int hc = 0;
  for (int l = 0; l < L && !rc; ++l) {
    const FermiLayer& ly = fl[l];
    rc = launch_pair_means(fd, H2[hc], ly.f2, Xc, ly.kmain, ly.f1, xbf, s);
    if (!rc) rc = launch_spin_means(fd, Xc, ly.kmain, ly.f1, g, ly.kg, xbf, s);
    if (rc) break;
    GemmArgs ga{};
    ga.X = g;
    ga.ldx = ly.kg;
    ga.Wt = ly.Wg;
    ga.Wt_lo = ly.Wg_lo;
    ga.wplane = (long)ly.d1 * ly.kg;
    ga.ldw = ly.kg;
    ga.Y = G;
    ga.ldy = ly.d1;
    ga.rows = Bc * RP;
    ga.K = ly.kg;
    ga.ncols = ly.d1;
    ga.R = 1;
    ga.lrp = lap ? RP : 0;
    ga.act = ACT_NONE;
    rc = launch_gemm(ga, prec, false, s);
    if (rc) return jq_fail(JQ_EINVAL, "spin-mean GEMM launch rejected");
    const int ldn = (l + 1 < L) ? fl[l + 1].kmain : korb;
    GemmArgs gm{};
    gm.X = Xc;
    gm.ldx = ly.kmain;
    gm.Wt = ly.Wm;
    gm.Wt_lo = ly.Wm_lo;
    gm.wplane = (long)ly.d1 * ly.kmain;
    gm.ldw = ly.kmain;
    gm.Y = Xn;
    gm.ldy = ldn;
    gm.rows = Bc * N * RP;
    gm.K = ly.kmain;
    gm.ncols = ly.d1;
    gm.bias = ly.b;
    gm.G = G;
    gm.ldg = ly.d1;
    gm.R = RP;
    gm.act = ACT_TANH;
    gm.Res = ly.res1 ? Xc : nullptr;
    gm.ldres = ly.kmain;
    gm.res_scale = 0.70710678118654752f;
    gm.Bc = Bc;
    gm.n_up = n_up;
    gm.n_dn = n_dn;
    {
      ProfScope ps(lap ? PROF_GEMM_LAYER_LAP : -1, s);
      rc = launch_gemm(gm, prec, xbf, s);
    }
    if (rc) return jq_fail(JQ_EINVAL, "layer GEMM launch rejected");
    if (ly.has_pair) {
      ProfScope ps(lap ? PROF_PAIR_LAP : -1, s);
      rc = launch_pair_layer(fd, H2[hc], ly.f2, H2[hc ^ 1], ly.d2, ly.W2, ly.b2, ly.res2, s);
      hc ^= 1;
    }
    std::swap(Xc, Xn);
  }
  if (rc) return jq_fail(JQ_EHIP, "FermiNet kernel launch failed");
  return orbitals_and_dets(elec, Xc, Bc, lap, s);
}

// Orbital projection per spin channel, envelope, determinants.
int jq_model::orbitals_and_dets(const float* elec, const void* Xlast, int Bc, bool lap, hipStream_t s) {
  const FermiDims fd = dims(Bc, lap);
  const int RP = fd.Rp;
  for (int c = 0; c < nch; ++c) {
    const long u0 = (nch == 1 || c == 0) ? 0 : (long)Bc * n_up;
    const long nu = nch == 1 ? (long)Bc * N : (c == 0 ? (long)Bc * n_up : (long)Bc * n_dn);
    GemmArgs go{};
    go.X = (const char*)Xlast + u0 * RP * korb * es();
    go.ldx = korb;
    go.Wt = Worb[c];
    go.Wt_lo = Worb_lo[c];
    go.wplane = (long)K * N * korb;
    go.ldw = korb;
    go.Y = Phi + u0 * RP * ldphi;
    go.ldy = ldphi;
    go.rows = (int)(nu * RP);
    go.K = korb;
    go.ncols = K * N;
    go.R = lap ? RP : 1;  // Laplacian mode: unit-tiled launch
    go.act = ACT_NONE;
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

int jq_model::forward(const float* elec, int b0, int Bc, bool lap, float* logpsi, float* sign, double* e_kin,
                      double* e_pot, double* e_loc, float* grad, hipStream_t s, bool lp_local) {
  const float* ec = elec + (long)b0 * T;
  ProfScope ps(lap ? PROF_LOCAL_ENERGY : PROF_FORWARD, s);
  int rc = kind == JQ_KIND_FERMINET ? forward_ferminet(ec, Bc, lap, s) : forward_psiformer(ec, Bc, lap, s);
  if (rc) return rc;
  if (skip_det) return 0;
  CombineArgs ca{};
  ca.det = det;
  ca.B_off = b0;
  ca.elec = elec;
  ca.atoms = atoms;
  ca.charges = charges;
  ca.jastrow = (kind == JQ_KIND_PSIFORMER && d.jastrow) ? 1 : 0;
  ca.alpha_par = jas[0];
  ca.alpha_anti = jas[1];
  ca.logpsi = logpsi;
  ca.sign = sign;
  ca.e_kin = e_kin;
  ca.e_pot = e_pot;
  ca.e_loc = e_loc;
  ca.grad = grad;
  ca.lapmode = lap ? 1 : 0;
  ca.lp_local = lp_local ? 1 : 0;
  launch_combine(dims(Bc, lap), ca, s);
  return 0;
}
