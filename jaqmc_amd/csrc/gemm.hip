// Host entry points of the GEMM family: argument checks and the dispatch to
// the per-mode instantiations (gemm_f32/bf16/x3/x6.hip, kernels in
// gemm_impl.h), plus a synthetic-buffer timing hook.

#include <vector>

#include "kernels.h"

int g_gemm_variant = 0;
int g_gemm_nw = 0;

int gemm_dispatch_f32(const GemmArgs& a, hipStream_t s);
int gemm_dispatch_bf16(const GemmArgs& a, hipStream_t s);
int gemm_dispatch_x3(const GemmArgs& a, hipStream_t s);
int gemm_dispatch_x6(const GemmArgs& a, hipStream_t s);
int gemm_dispatch_h3(const GemmArgs& a, hipStream_t s);
void bgemm_dispatch_f32(const BGemmArgs& a, hipStream_t s);
void bgemm_dispatch_bf16(const BGemmArgs& a, hipStream_t s);
void bgemm_dispatch_x3(const BGemmArgs& a, hipStream_t s);
void bgemm_dispatch_x6(const BGemmArgs& a, hipStream_t s);

int launch_bgemm(const BGemmArgs& a, int prec, hipStream_t s) {
  if (a.K % 32 != 0 || a.ldx % 4 != 0 || (!a.nn && a.ldw % 4 != 0) || (a.nn && (a.ldw % 4 || a.ncols % 4)))
    return -1;
  if (a.rows <= 0 || a.ncols <= 0 || a.nw * a.nh <= 0) return 0;
  switch (prec) {
    case PREC_F32: bgemm_dispatch_f32(a, s); return 0;
    case PREC_BF16: bgemm_dispatch_bf16(a, s); return 0;
    case PREC_BF16X3: bgemm_dispatch_x3(a, s); return 0;
    case PREC_BF16X6:
    case PREC_FP16X3: bgemm_dispatch_x6(a, s); return 0;
  }
  return -1;
}

// Supported Laplacian-mode unit heights (rows per unit after padding).
int gemm_unit_rows(int R) {
  const int opts[] = {16, 32, 48, 64, 96, 128, 192};
  for (int o : opts)
    if (R <= o) return o;
  return -1;
}

// The tile table gives 128-row units in FP16X3 one unit per tile and four
// 16-column blocks per wave (256 columns): a workgroup then holds every
// column of its unit.  The dispatch rejects an Yln launch of any other tile.
bool gemm_ln_fusable(int R, int ncols, int prec) {
  return R == 128 && ncols == 256 && prec == PREC_FP16X3 && g_gemm_variant == 0;
}

int launch_gemm(const GemmArgs& a, int prec, bool y_bf16, hipStream_t s) {
  if (y_bf16) return -1;
  if (a.K % 32 != 0 || a.ldx % 4 != 0 || a.ldw % 4 != 0) return -1;
  if (a.R > 1 && a.rows % a.R != 0) return -1;
  if (a.rgrp > 0 && (a.R > 1 || a.Res || a.G || a.hs_H > 0 || a.nbatch > 1 || a.lrp > 0)) return -1;
  if (a.Yln && (a.R <= 1 || a.G || a.hs_H > 0 || a.nbatch > 1 || a.ldy != a.ncols || a.ln_R > a.R ||
                a.ln_R < ROW_J0 || !a.ln_s || !a.ln_b || (a.Res && a.ldres % 4 != 0)))
    return -1;
  if (a.rows == 0) return 0;
  switch (prec) {
    case PREC_F32: return gemm_dispatch_f32(a, s);
    case PREC_BF16: return gemm_dispatch_bf16(a, s);
    case PREC_BF16X3: return gemm_dispatch_x3(a, s);
    case PREC_BF16X6: return gemm_dispatch_x6(a, s);
    // value-only rows keep the BF16X6 arithmetic (no per-unit row scales)
    case PREC_FP16X3: return (a.R <= 1 && a.lrp <= 0) ? gemm_dispatch_x6(a, s) : gemm_dispatch_h3(a, s);
  }
  return -1;
}

// Benchmark hook: time `reps` launches of one GEMM configuration on synthetic
// device buffers (HIP events on the null stream).  variant % 100 as
// g_gemm_variant; variant % 1000 >= 100 selects the plain row-tile-major
// order; variant / 1000 > 0 sets the tangent-unit rows of plain (R = 1)
// launches (48 otherwise).
extern "C" int jq_bench_gemm(int prec, int R, int rows, int K, int ncols, int act, int variant, int reps,
                             float* ms_out) {
  if (rows <= 0 || K <= 0 || ncols <= 0 || K % 32 || (R > 1 && rows % R) || reps <= 0 || prec < 0 ||
      prec > PREC_FP16X3)
    return -1;
  // weight planes: NS BF16 planes; FP16X3 adds two FP16 planes and 2*ncols
  // FP32 column scales after three BF16 planes
  const size_t wbytes = prec == PREC_FP16X3 ? (size_t)ncols * K * 5 * 2 + (size_t)ncols * 8
                                            : (size_t)ncols * K * 2 * (prec == PREC_BF16X6 ? 3 : (prec == PREC_BF16X3 ? 2 : 1));
  // G holds one (R x ncols) block per walker of the unit -> walker map below
  // (7 + 7 electrons): unit u belongs to a walker below units / 7 + 1
  const long units = R > 1 ? rows / R : rows;
  const size_t gbytes = (size_t)ncols * 4 * (R > 1 ? R : 1) * (units / 7 + 2);
  float *X = nullptr, *W = nullptr, *Y = nullptr, *G = nullptr, *b = nullptr;
  bf16* Wb = nullptr;
  const long nW = (long)ncols * K;
  int rc = 0;
  if (hipMalloc(&X, (size_t)rows * K * 4) || hipMalloc(&W, nW * 4) || hipMalloc(&Y, (size_t)rows * ncols * 4) ||
      hipMalloc(&Wb, wbytes) || hipMalloc(&G, gbytes) || hipMalloc(&b, ncols * 4))
    rc = -3;
  if (!rc) {
    (void)hipMemset(X, 0, (size_t)rows * K * 4);
    (void)hipMemset(W, 0, nW * 4);
    (void)hipMemset(Wb, 0, wbytes);
    (void)hipMemset(b, 0, ncols * 4);
    (void)hipMemset(G, 0, gbytes);
    if (prec == PREC_FP16X3) {  // unit column scales
      std::vector<float> ones(2 * (size_t)ncols, 1.f);
      (void)hipMemcpy((char*)Wb + (size_t)ncols * K * 10, ones.data(), ones.size() * 4, hipMemcpyHostToDevice);
    }
    GemmArgs a{};
    a.X = X;
    a.ldx = K;
    a.Wt = W;
    a.ldw = K;
    a.Wt_lo = Wb;
    a.wplane = nW;
    a.Y = Y;
    a.ldy = ncols;
    a.rows = rows;
    a.K = K;
    a.ncols = ncols;
    a.bias = b;
    a.R = R;
    a.act = act ? ACT_TANH : ACT_NONE;
    a.res_scale = 1.f;
    if (act == 2 && ncols <= K) {  // realistic layer epilogue: G broadcast + residual
      a.Bc = units / 14 > 0 ? (int)(units / 14) : 1;
      a.n_up = 7;
      a.n_dn = 7;
      a.G = G;
      a.ldg = ncols;
      a.Res = X;
      a.ldres = K;
    }
    if (R <= 1) a.lrp = variant % 1000000 >= 1000 ? variant % 1000000 / 1000 : 48;  // plain rows as lrp-row tangent units
    a.linear_tiles = variant % 1000 >= 100 ? 1 : 0;
    const int old = g_gemm_variant, old_nw = g_gemm_nw;
    g_gemm_variant = variant % 100;
    g_gemm_nw = variant >= 1000000 ? 4 : 0;  // variant + 1000000: four-wave tiles only
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    if (!rc) rc = launch_gemm(a, prec, false, 0);
    (void)hipEventRecord(e0, 0);
    for (int i = 0; i < reps && !rc; ++i) rc = launch_gemm(a, prec, false, 0);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_out = ms / reps;
    g_gemm_variant = old;
    g_gemm_nw = old_nw;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(X);
  (void)hipFree(W);
  (void)hipFree(Y);
  (void)hipFree(Wb);
  (void)hipFree(G);
  (void)hipFree(b);
  return rc;
}
