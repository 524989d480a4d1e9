"""CPU oracle for the per-walker VMC hot path.

TEST INFRASTRUCTURE.  A numpy restatement of the reference's algorithm
(FermiNet / Psiformer log|psi|, dense forward Laplacian, slogdet, Coulomb
potential, Metropolis step, walker reductions), evaluated in float64 (the
FP64 reference) or float32 (the plain-FP32 reference).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker; the product path (``jaqmc_amd``) never
imports it and fails loudly when its HIP library is missing.

How each function is pinned (checks that a plausible mistake would fail):

* ``lap.Lap`` rules and the whole ``wavefunction.logpsi`` Laplacian: against a
  brute-force float64 autograd Hessian trace (torch double backward) on small
  systems, and against central finite differences -- tests/test_oracle.py.
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

Parity unpinned: flax's LayerNorm / MultiHeadDotProductAttention internals
(variance formula, scaling placement) are restated from flax's published
definitions; no reference fixture pins them.  The JAX threefry PRNG stream is
not reproduced: parity runs feed explicit normals and uniforms to both sides.
"""

from . import hamiltonian, lap, wavefunction  # noqa: F401
