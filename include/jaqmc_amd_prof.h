/*
 * jaqmc_amd measurement hooks.  No counterpart in the reference (which has no
 * profiler integration, only a wall-clock TimeTracker, utils/time_tracker.py);
 * these serve benchmark harnesses that need per-kernel-class device time.
 *
 * When enabled, every launch of an instrumented kernel class is bracketed by
 * two HIP events recorded on the stream the kernel is launched on.
 */
#ifndef JAQMC_AMD_PROF_H
#define JAQMC_AMD_PROF_H
#ifdef __cplusplus
extern "C" {
#endif

#define JQ_PROF_GEMM_LAYER_LAP 0 /* one-electron layer GEMMs, Laplacian mode */
#define JQ_PROF_DET_LAP 1        /* determinant + Laplacian kernel          */
#define JQ_PROF_GEMM_ORB_LAP 2   /* orbital projection GEMMs, Laplacian mode */
#define JQ_PROF_PAIR_LAP 3       /* two-electron stream, Laplacian mode      */
#define JQ_PROF_FORWARD 4        /* whole log|psi| chunk                     */
#define JQ_PROF_LOCAL_ENERGY 5   /* whole local-energy chunk                 */
#define JQ_PROF_ATTN_LAP 6       /* Psiformer attention, Laplacian mode      */

/* Enable (1, also clears the record) or disable (0) recording. */
int jq_profile_enable(int on);
/* Total device milliseconds and launch count of one class since enabling
 * (synchronises on the recorded events). */
int jq_profile_read(int cls, double* total_ms, int* count);

/* Time `reps` launches of the tangent GEMM kernel on synthetic zero buffers:
 * Y (rows x ncols) = X (rows x K) W^T, R tangent rows per unit (1: plain rows),
 * act 0/1 (none/tanh rule), prec JQ_PREC_*, variant 0 (default tiling) or
 * 10*units_per_tile + col_blocks_per_wave.  Average milliseconds -> *ms. */
int jq_bench_gemm(int prec, int R, int rows, int K, int ncols, int act, int variant, int reps, float* ms);

#ifdef __cplusplus
}
#endif
#endif
