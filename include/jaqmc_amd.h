/*
 * jaqmc_amd -- C ABI of the MI355X (gfx950) per-walker VMC hot path.
 *
 * One shared library, libjaqmc_amd.so.  Every array argument is a DEVICE
 * pointer, row-major, walker-major; every call is enqueued on the given HIP
 * stream with no hidden host synchronisation.  Calls return 0 on success or a
 * negative JQ_E* code; jq_last_error() gives the message of the calling
 * thread's last failure.  A non-finite local energy is data, not an error (the
 * reference reduces with nanmean, estimator/base.py:44).
 *
 * Threading: calls on distinct jq_model handles may run concurrently on
 * distinct streams; one handle must not be used from two streams at once (its
 * workspace is shared).  Parameters are immutable after jq_model_create.
 *
 * The reference is Python/JAX with no native FFI of its own; each entry point
 * below replaces one reference interface, cited per function (paths relative
 * to the reference's src/jaqmc/).
 */
#ifndef JAQMC_AMD_H
#define JAQMC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JQ_OK 0
#define JQ_EINVAL -1   /* bad argument / unsupported shape */
#define JQ_EHIP -2     /* a HIP runtime call failed */
#define JQ_ENOMEM -3   /* device allocation failed */

#define JQ_KIND_FERMINET 0
#define JQ_KIND_PSIFORMER 1

#define JQ_ENV_ISOTROPIC 0
#define JQ_ENV_ABS_ISOTROPIC 1
#define JQ_ENV_NULL 2
#define JQ_ENV_DIAGONAL 3      /* EnvelopeType.diagonal: rejected (JQ_EINVAL)        */

/* PsiformerLayer.layer_norm_mode (wavefunction/backbone/psiformer.py:13-26,70-98) */
#define JQ_LN_PRE 0
#define JQ_LN_POST 1
#define JQ_LN_NULL 2

/* jq_model_desc.options: module options of the reference that change the
 * computation.  A pack always carries every bias slot of the default module;
 * a module built without a bias (with_bias / input_bias false) passes zeros
 * there and sets the flag, and create rejects a flag whose slots are not all
 * zero.  Unsupported options return JQ_EINVAL before any allocation. */
#define JQ_OPT_NO_ATTN_BIAS (1 << 0)   /* Psiformer with_bias=False (psiformer.py:57,80)            */
#define JQ_OPT_NO_INPUT_BIAS (1 << 1)  /* Psiformer input_bias=False (backbone/psiformer.py:175)    */
#define JQ_OPT_ORBITAL_BIAS (1 << 2)   /* bias_orbitals=True: pack holds one (K*N) bias per orbital
                                          kernel (output/orbital.py:60-78)                         */
#define JQ_OPT_NO_RESCALE (1 << 3)     /* Psiformer rescale=False: [r, d] features (input/atomic.py) */
#define JQ_OPT_SHARED_ORBITALS (1 << 4)/* orbitals_spin_split=False: one orbital kernel and one
                                          envelope for both channels (orbital.py:73-76)            */
#define JQ_OPT_USE_LAST_LAYER (1 << 5) /* FermiNet use_last_layer=True: rejected (JQ_EINVAL)      */
/* Library option (no reference counterpart): allocate the parameter-gradient
 * workspace at create, for jq_logpsi_vjp (FermiNet with hidden widths
 * multiples of 32; Psiformer with N <= 64 and heads_dim 32 or 64). */
#define JQ_OPT_WITH_GRADIENTS (1 << 16)

/* Arithmetic of the layer / tangent GEMMs.  Determinants (LU, inverse,
 * Laplacian traces) and walker reductions are FP64 in every mode. */
#define JQ_PREC_FP32 0   /* FP32 MFMA                                               */
#define JQ_PREC_BF16 1   /* BF16 operands (one term), FP32 accumulate              */
#define JQ_PREC_BF16X3 2 /* two-term split BF16 (hh + hl + lh), FP32 accumulate    */
#define JQ_PREC_BF16X6 3 /* three-term split BF16 (6 products), FP32 accumulate:
                            FP32-level accuracy on BF16 matrix cores              */
#define JQ_PREC_FP16X3 4 /* two-term split FP16 (hh + hl + lh), FP32 accumulate:
                            2^-22 operand error on 16-bit matrix cores; each
                            unit's Laplacian row is scaled by a power of two
                            before the split so it stays in FP16 range        */
/* Activations are FP32 in HBM in every mode; operands are split into 16-bit
 * terms when staged into LDS, weights are uploaded pre-split. */

#define JQ_MAX_LAYERS 8

typedef struct jq_model_desc {
  int kind;              /* JQ_KIND_*                                            */
  int n_up, n_dn;        /* electrons per spin channel (ordered up then down)   */
  int n_atoms;
  int ndets;
  int n_layers;
  /* FermiNet: one- and two-electron stream widths per layer */
  int hidden_single[JQ_MAX_LAYERS];
  int hidden_double[JQ_MAX_LAYERS];
  /* Psiformer */
  int num_heads, heads_dim;
  int n_mlp;             /* number of hidden MLP layers (reference default 1)    */
  int mlp_hidden[4];
  int jastrow;           /* 1: SimpleEEJastrow                                   */
  int envelope;          /* JQ_ENV_*                                             */
  int precision;         /* JQ_PREC_*                                            */
  int chunk_walkers;     /* walkers per workspace chunk (0: library default)     */
  const float* atoms;    /* HOST (n_atoms, 3); copied at create                  */
  const float* charges;  /* HOST (n_atoms,);  copied at create                   */
  int layer_norm_mode;   /* JQ_LN_* (Psiformer; 0 = the reference default)       */
  int options;           /* JQ_OPT_* bits (0 = the reference defaults)           */
} jq_model_desc;

typedef struct jq_model jq_model;

/* Create a model from the flat parameter pack (HOST float32 buffer in the
 * order documented in jaqmc_amd/params.py).  Replaces Wavefunction.init_params
 * + the Flax parameter tree consumed by MoleculeWavefunction.evaluate
 * (wavefunction/base.py:45-53; app/molecule/wavefunction/ferminet.py:44-74,
 * app/molecule/wavefunction/psiformer.py:93-134). */
int jq_model_create(const jq_model_desc* desc, const float* packed_params, size_t n_floats,
                    jq_model** out);
/* Destroy: waits only for the streams this model was used on (events recorded
 * at each call) and releases its memory stream-ordered; no device-wide sync. */
int jq_model_destroy(jq_model* m);
/* Bytes of device workspace the model holds (weights + chunk buffers). */
size_t jq_model_workspace_bytes(const jq_model* m);

/* log|psi| and sign for B walkers.  Replaces vmap(MoleculeWavefunction.logpsi)
 * / phase_logpsi (app/molecule/wavefunction/ferminet.py:98-127,
 * app/molecule/wavefunction/psiformer.py:168-197). elec: (B, N, 3). */
int jq_logpsi(jq_model* m, const float* elec, int B, float* logpsi, float* sign, void* stream);

/* n_steps all-electron Metropolis moves.  Replaces the fori_loop of
 * MCMCSampler._mh_update (sampler/mcmc.py:96-137,173-178) with
 * batch_log_prob = 2*logpsi (sampler/base.py:174).
 *   elec    (B, N, 3)  in/out
 *   logp    (B,)       in/out: 2*log|psi| at elec (jq_logpsi output doubled)
 *   normals (n_steps, B, N, 3) or NULL -> Philox-4x32-10 normals keyed by
 *           (seed, ctr + step, walker, element)
 *   uniforms (n_steps, B) or NULL -> Philox uniforms on the same key
 *   n_accept (B,) int32, accumulated (not cleared) accept counts
 * Accept iff 2log|psi(x')| - logp > log(u) (strict). */
int jq_mh_steps(jq_model* m, float* elec, float* logp, int B, int n_steps, float width,
                const float* normals, const float* uniforms, uint64_t seed, uint64_t ctr,
                int32_t* n_accept, void* stream);

/* Per-walker local energy E_L = -1/2 (lap log|psi| + |grad log|psi||^2) + V by
 * the forward Laplacian.  Replaces EuclideanKinetic(mode=forward_laplacian)
 * (estimator/kinetic/euclidean.py:114-135, _common.py:61-73), potential_energy
 * (app/molecule/hamiltonian.py:9-22) and TotalEnergy (estimator/total_energy.py:36-60).
 *   e_kin, e_pot, e_loc (B,) float64;  logpsi (B,) float32 or NULL;
 *   grad (B, 3N) float32 or NULL. */
int jq_local_energy(jq_model* m, const float* elec, int B, double* e_kin, double* e_pot,
                    double* e_loc, float* logpsi, float* grad, void* stream);

/* Orbital matrices A[w, k, i, o] = Phi[k, i, o] * env[k, i, o] (FP32), the
 * (B, K, N, N) array of MoleculeWavefunction.orbitals / get_orbitals
 * (app/molecule/wavefunction/ferminet.py:80-96,128-138;
 * app/molecule/wavefunction/psiformer.py:160-166,202-212).  Row = electron,
 * column = orbital. */
int jq_orbitals(jq_model* m, const float* elec, int B, float* orbitals, void* stream);

/* Parameter gradient of log|psi| contracted with per-walker weights:
 *   grad = sum_w cot[w] d log|psi(x_w)| / d theta      (FP32, pack order)
 * the per-walker value_and_grad of LossAndGrad.evaluate_single_walker
 * (estimator/loss_grad.py:70-103) reduced with the walker weights of
 * finalize_stats (:115-128; jq_loss_clip's weights give the VMC energy
 * gradient).  cot (B,) float64 on the device; a walker with cot 0 contributes
 * exactly 0 (its determinants are not inverted).  FermiNet and Psiformer
 * models created with JQ_OPT_WITH_GRADIENTS; FP32 arithmetic, FP64
 * determinant adjoint and FP64 accumulation over walkers. */
int jq_logpsi_vjp(jq_model* m, const float* elec, int B, const double* cot, float* grad, void* stream);
/* Floats in the model's parameter pack (the length of jq_logpsi_vjp's grad). */
size_t jq_model_param_count(const jq_model* m);

/* Device-resident sampler state of jq_sample (MCMCState, sampler/mcmc.py:39-50):
 * double[3 + adapt_frequency] = {stddev, counter, accepted-move total,
 * pmoves[adapt_frequency]}.  Initialise to {width, 0, 0, 0...}. */
#define JQ_MCMC_STATE_LEN(af) (3 + (af))

/* n_outer sampler steps of MCMCSampler.step (sampler/mcmc.py:139-197), all on
 * the device with no host synchronisation: each recomputes logp = 2 log|psi|
 * at the current walkers, makes `steps` all-electron moves with the width in
 * state[0], then updates the state exactly as the reference does (counter += 1,
 * pmoves[counter % af] = pmove, width x/÷ 1.1 when counter % af == 0 and the
 * mean pmove leaves [pmove_lo, pmove_hi]).  With n_outer = burn_in this is the
 * burn-in loop of SamplingStage (workflow/stage/sampling.py:246-301).
 *   elec (B, N, 3) in/out;  logp (B,) out: 2 log|psi| at the final walkers;
 *   n_accept (B,) int32 accumulated; state[2] must equal its sum on entry
 *   (zeroed counts with state[2] = 0);  pmove_out (n_outer,) float64 or NULL.
 * Philox counters: outer step s, move t uses ctr + s * steps + t. */
int jq_sample(jq_model* m, float* elec, float* logp, int B, int n_outer, int steps, double* state,
              int adapt_frequency, double pmove_lo, double pmove_hi, uint64_t seed, uint64_t ctr,
              int32_t* n_accept, double* pmove_out, void* stream);

/* Mean and E[x^2]-E[x]^2 over the finite entries of x (B,) in FP64, written to
 * out[0], out[1]; out[2] = count of finite entries.  Replaces mean_reduce
 * (estimator/base.py:27-53) on one device (nanmean semantics). */
int jq_reduce_mean_var(const double* x, int B, double* out, void* stream);

/* Sum of int32 accept counts (B,) into out[0] (float64).  The pmove numerator
 * of MCMCSampler.step (sampler/mcmc.py:179-180). */
int jq_reduce_accepts(const int32_t* n_accept, int B, double* out, void* stream);

/* LossAndGrad's loss clipping and gradient weights (estimator/loss_grad.py:
 * 91-128, utils/clip.py:11-60), FP64 on the device:
 *   method 0 "mad":  clip to median -+ scale * median(|x - median|)
 *   method 1 "iqr":  clip to [q1 - scale iqr, q3 + scale iqr]
 *   method 2 "none"
 * with NaN-ignoring numpy-"linear" quantiles.  Writes clipped (B,) (or NULL),
 * weights (B,) = 2 (clipped - mean(clipped)) / n over the n non-NaN walkers
 * (0 for NaN walkers), so sum_w weights_w grad log|psi_w| is the reference's
 * loss gradient, and stats[5] = {nanmean(x), nanmean(clipped), lo, hi, n}
 * (or NULL).  workspace: device memory of jq_loss_clip_workspace_bytes(B)
 * bytes (partial histograms of the radix select; contents need not be zero). */
#define JQ_LOSS_CLIP_MAX_B (1 << 28)
size_t jq_loss_clip_workspace_bytes(int B);
int jq_loss_clip(const double* x, int B, int method, double scale, double* clipped, double* weights,
                 double* stats, void* workspace, size_t ws_bytes, void* stream);

/* jq_loss_clip without a workspace argument.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops or harnesses that forbid hipMalloc).  Same kernels and
 * results as jq_loss_clip; the scratch is a static device array of
 * JQ_LOSS_CLIP_SCRATCH_BYTES per GPU (one copy per device, held until the
 * process ends), so the call allocates nothing.  B above
 * JQ_LOSS_CLIP_NOALLOC_MAX_B is rejected with JQ_EINVAL.  All calls on one GPU
 * share that array and every call rewrites it before reading it: calls that
 * overlap on one GPU (two streams or two host threads at once) race. */
#define JQ_LOSS_CLIP_NOALLOC_MAX_B (1 << 28)
#define JQ_LOSS_CLIP_SCRATCH_BYTES 262464
int jq_loss_clip_noalloc(const double* x, int B, int method, double scale, double* clipped, double* weights,
                         double* stats, void* stream);

const char* jq_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* JAQMC_AMD_H */
