"""CPU oracle for the per-walker VMC hot path.

TEST INFRASTRUCTURE.  A numpy restatement of the reference's algorithm
(FermiNet / Psiformer log|psi|, dense forward Laplacian, slogdet, Coulomb
potential, Metropolis step, walker reductions), evaluated in float64 (the
FP64 reference) or float32 (the plain-FP32 reference).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker; the product path (``jaqmc_amd``) never
imports it and fails loudly when its HIP library is missing.

How each function is pinned (checks that a plausible mistake would fail):

* ``lap.Lap`` rules and the whole ``wavefunction.logpsi`` pipeline (FermiNet
  and Psiformer, every LayerNorm placement and module option): value, gradient
  and Laplacian against an independent float64 torch evaluation built from
  torch library operations (layer_norm, softmax, linalg.slogdet) with autograd
  gradients and the autograd Hessian trace (double backward) --
  tests/test_oracle_pins.py; and against central finite differences --
  tests/test_oracle.py.
* ``layer_norm``: zero mean and mean square v/(v+eps) per row, a variance-eps
  row (eps inside the rsqrt), a constant row (gives the bias), torch's
  layer_norm.
* ``mha``: uniform weights when scores vanish, one electron, a hand-computed
  two-electron softmax (pins the 1/sqrt(dh) placement), torch's
  scaled_dot_product_attention.
* ``aggregate``: hand values of an asymmetric two-electron stream (mean over
  the first index), two channels and polarized.
* kinetic machinery: closed forms -- hydrogen-like logpsi = -r gives
  E_L = -1/2 exactly; harmonic logpsi = -c sum x^2 gives KE = c*n*D - 2c^2 sum x^2
  (reference tests: tests/hydrogen/atom_test.py:31-53,
  tests/estimator/kinetic_forward_laplacian_test.py:42-86).
* slogdet: numpy.linalg.slogdet / brute-force permutation expansion (n <= 4).
* antisymmetry: same-spin swap flips the sign and keeps |psi|
  (tests/wavefunction/molecule_wavefunction_test.py:67-121).
* Jastrow cusp slopes 1/4 and 1/2 (tests/wavefunction/jastrow_test.py:21-100).
* potential: brute-force double loops; the 3-4-5 triangle
  (tests/geometry/geometry_test.py:20-79).
* Metropolis: the exact rule on given (xi, u); stationary moments of a known
  Gaussian |psi|^2.
* reductions: float64 two-pass vs numpy.

Parity unpinned: only what no check here can reach -- the JAX threefry PRNG
stream (parity runs feed explicit normals and uniforms to both sides instead)
and flax's float32 summation order inside LayerNorm's variance (the oracle
runs in float64, where it does not matter).
"""

from . import hamiltonian, lap, wavefunction  # noqa: F401
