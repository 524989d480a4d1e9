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

/* Arithmetic of the layer / tangent GEMMs.  Determinants (LU, inverse,
 * Laplacian traces) and walker reductions are FP64 in every mode. */
#define JQ_PREC_FP32 0   /* FP32 MFMA                                               */
#define JQ_PREC_BF16 1   /* BF16 operands (one term), FP32 accumulate              */
#define JQ_PREC_BF16X3 2 /* two-term split BF16 (hh + hl + lh), FP32 accumulate    */
#define JQ_PREC_BF16X6 3 /* three-term split BF16 (6 products), FP32 accumulate:
                            FP32-level accuracy on BF16 matrix cores              */
/* Activations are FP32 in HBM in every mode; operands are split into BF16
 * terms when staged into LDS. */

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
} jq_model_desc;

typedef struct jq_model jq_model;

/* Create a model from the flat parameter pack (HOST float32 buffer in the
 * order documented in jaqmc_amd/params.py).  Replaces Wavefunction.init_params
 * + the Flax parameter tree consumed by MoleculeWavefunction.evaluate
 * (wavefunction/base.py:45-53; app/molecule/wavefunction/ferminet.py:44-74,
 * app/molecule/wavefunction/psiformer.py:93-134). */
int jq_model_create(const jq_model_desc* desc, const float* packed_params, size_t n_floats,
                    jq_model** out);
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

/* Mean and E[x^2]-E[x]^2 over the finite entries of x (B,) in FP64, written to
 * out[0], out[1]; out[2] = count of finite entries.  Replaces mean_reduce
 * (estimator/base.py:27-53) on one device (nanmean semantics). */
int jq_reduce_mean_var(const double* x, int B, double* out, void* stream);

/* Sum of int32 accept counts (B,) into out[0] (float64).  The pmove numerator
 * of MCMCSampler.step (sampler/mcmc.py:179-180). */
int jq_reduce_accepts(const int32_t* n_accept, int B, double* out, void* stream);

const char* jq_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* JAQMC_AMD_H */
