"""Dense forward-Laplacian arrays for the CPU oracle (numpy).

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP path: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path never routes through it.

A :class:`Lap` carries, for a batch of walkers, the value ``v`` of a quantity,
its Jacobian ``J`` with respect to all ``T = 3N`` electron coordinates and its
Laplacian ``lap`` (the trace of the Hessian over the same coordinates):

    v   : (B, *S)
    J   : (B, T, *S)
    lap : (B, *S)

The propagation rules restate the reference's forward-Laplacian rules with a
*dense* Jacobian (the reference additionally exploits Local1/Local2 sparsity,
which changes cost, not values):

* linear maps (``dot_general``): J <- J W, lap <- lap W
  (laplacian/primitives/dot_general.py:377-407);
* elementwise ``f``: J <- f'(x) J, lap <- f'(x) lap + f''(x) sum_t J_t^2
  (laplacian/primitives/elementwise.py:42-72);
* product: lap(xy) = y lap x + x lap y + 2 sum_t J_x,t J_y,t
  (laplacian/primitives/core.py:450-501);
* reductions over electrons / features are linear (reductions.py:25-61);
* slogdet: grad = tr(A^-1 J_t), lap = tr(A^-1 L) - sum_t tr((A^-1 J_t)^2)
  (laplacian/primitives/slogdet.py:44-59).

With ``T = 0`` the same code evaluates plain values (no tangents), which is how
the oracle computes log|psi| for the Metropolis step.
"""

from __future__ import annotations

import numpy as np


class Lap:
    __slots__ = ("v", "J", "lap")

    def __init__(self, v, J, lap):
        self.v = v
        self.J = J
        self.lap = lap

    # -- construction -----------------------------------------------------
    @staticmethod
    def const(v, B, T, dtype):
        """A quantity with no coordinate dependence; v has no batch axis."""
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=dtype), (B,) + np.shape(v)))
        return Lap(v, np.zeros((B, T) + v.shape[1:], dtype), np.zeros_like(v))

    @staticmethod
    def coords(x, with_tangents=True):
        """Seed: electron coordinates x (B, N, 3) with J = identity over 3N."""
        B, N, _ = x.shape
        T = 3 * N if with_tangents else 0
        J = np.zeros((B, T, N, 3), x.dtype)
        if with_tangents:
            eye = np.eye(3 * N, dtype=x.dtype).reshape(3 * N, N, 3)
            J[:] = eye[None]
        return Lap(x.copy(), J, np.zeros_like(x))

    @property
    def shape(self):
        return self.v.shape

    @property
    def T(self):
        return self.J.shape[1]

    @property
    def dtype(self):
        return self.v.dtype

    def _jaxis(self, axis):
        """Map an axis of v (batch-first) to the matching axis of J."""
        nd = self.v.ndim
        if axis < 0:
            axis += nd
        return axis if axis == 0 else axis + 1

    # -- linear structure ---------------------------------------------------
    def __add__(self, o):
        if isinstance(o, Lap):
            v = self.v + o.v
            B, T = v.shape[0], self.T
            return Lap(v, np.broadcast_to(self.J + o.J, (B, T) + v.shape[1:]),
                       np.broadcast_to(self.lap + o.lap, v.shape))
        v = self.v + np.asarray(o, dtype=self.dtype)
        B, T = v.shape[0], self.T
        return Lap(v, np.broadcast_to(self.J, (B, T) + v.shape[1:]), np.broadcast_to(self.lap, v.shape))

    __radd__ = __add__

    def __neg__(self):
        return Lap(-self.v, -self.J, -self.lap)

    def __sub__(self, o):
        return self + (-o)

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if isinstance(o, Lap):
            # product rule (laplacian/primitives/core.py:494-500)
            v = self.v * o.v
            J = self.J * _ex(o.v) + o.J * _ex(self.v)
            lap = self.lap * o.v + o.lap * self.v + 2.0 * np.sum(self.J * o.J, axis=1)
            return Lap(v, J, lap)
        o = np.asarray(o, dtype=self.dtype)
        v = self.v * o
        B, T = v.shape[0], self.T
        return Lap(v, np.broadcast_to(self.J * _exc(o, self.v.ndim), (B, T) + v.shape[1:]),
                   np.broadcast_to(self.lap * o, v.shape))

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, Lap):
            return self * o.recip()
        return self * (1.0 / np.asarray(o, dtype=self.dtype))

    def matmul(self, W, b=None):
        """Dense layer on the last axis: y = x W (+ b on the value only)."""
        W = np.asarray(W, dtype=self.dtype)
        v = self.v @ W
        if b is not None:
            v = v + np.asarray(b, dtype=self.dtype)
        return Lap(v, self.J @ W, self.lap @ W)

    def sum(self, axis):
        return Lap(self.v.sum(axis), self.J.sum(self._jaxis(axis)), self.lap.sum(axis))

    def mean(self, axis):
        n = self.v.shape[axis]
        return self.sum(axis) * (1.0 / n)

    def __getitem__(self, idx):
        """Index the non-batch axes: ``x[idx]`` applies to v[:, idx]."""
        if not isinstance(idx, tuple):
            idx = (idx,)
        return Lap(self.v[(slice(None),) + idx], self.J[(slice(None), slice(None)) + idx],
                   self.lap[(slice(None),) + idx])

    def reshape(self, *shape):
        B, T = self.v.shape[0], self.T
        return Lap(self.v.reshape((B,) + shape), self.J.reshape((B, T) + shape), self.lap.reshape((B,) + shape))

    def transpose(self, *axes):
        """Permute the non-batch axes (axes are 0-based over those axes)."""
        va = (0,) + tuple(a + 1 for a in axes)
        ja = (0, 1) + tuple(a + 2 for a in axes)
        return Lap(self.v.transpose(va), self.J.transpose(ja), self.lap.transpose(va))

    def expand(self, axis):
        """Insert a singleton axis (axis counted over v, batch-first)."""
        if axis < 0:
            axis += self.v.ndim + 1
        return Lap(np.expand_dims(self.v, axis), np.expand_dims(self.J, self._jaxis(axis)),
                   np.expand_dims(self.lap, axis))

    def broadcast_to(self, shape):
        B, T = self.v.shape[0], self.T
        return Lap(np.broadcast_to(self.v, (B,) + shape), np.broadcast_to(self.J, (B, T) + shape),
                   np.broadcast_to(self.lap, (B,) + shape))

    # -- elementwise --------------------------------------------------------
    def _elementwise(self, f, f1, f2):
        sq = np.sum(self.J * self.J, axis=1)
        return Lap(f, self.J * _ex(f1), f1 * self.lap + f2 * sq)

    def tanh(self):
        t = np.tanh(self.v)
        d1 = 1.0 - t * t
        return self._elementwise(t, d1, -2.0 * t * d1)

    def exp(self):
        e = np.exp(self.v)
        return self._elementwise(e, e, e)

    def log(self):
        return self._elementwise(np.log(self.v), 1.0 / self.v, -1.0 / (self.v * self.v))

    def log1p(self):
        u = 1.0 + self.v
        return self._elementwise(np.log(u), 1.0 / u, -1.0 / (u * u))

    def sqrt(self):
        s = np.sqrt(self.v)
        return self._elementwise(s, 0.5 / s, -0.25 / (s * s * s))

    def rsqrt(self):
        s = 1.0 / np.sqrt(self.v)
        return self._elementwise(s, -0.5 * s * s * s, 0.75 * s * s * s * s * s)

    def recip(self):
        r = 1.0 / self.v
        return self._elementwise(r, -r * r, 2.0 * r * r * r)

    def abs(self):
        s = np.sign(self.v)
        return self._elementwise(np.abs(self.v), s, np.zeros_like(s))

    def square(self):
        return self._elementwise(self.v * self.v, 2.0 * self.v, np.full_like(self.v, 2.0))


def concat(xs, axis):
    """Concatenate along a non-batch axis given batch-first (axis >= 1)."""
    if axis < 0:
        axis += xs[0].v.ndim
    return Lap(np.concatenate([x.v for x in xs], axis), np.concatenate([x.J for x in xs], axis + 1),
               np.concatenate([x.lap for x in xs], axis))


def _ex(v):
    """Insert the tangent axis into a value-shaped array for broadcasting with J."""
    return np.expand_dims(v, 1)


def _exc(c, vnd):
    """Broadcast a constant (shape broadcastable to v's non-batch axes) against J."""
    c = np.asarray(c)
    if c.ndim == vnd:  # includes the batch axis
        return np.expand_dims(c, 1)
    return c


def slogdet(A: Lap):
    """Forward Laplacian of slogdet over the last two axes.

    Restates laplacian/primitives/slogdet.py:44-59 in the oracle's dtype.
    Returns (sign, logdet Lap) with logdet shape = A.shape[:-2].
    """
    Ainv = np.linalg.inv(A.v)
    sign, ld = np.linalg.slogdet(A.v)
    # B_t = A^-1 J_t : (B, T, ..., n, n)
    Bt = np.matmul(np.expand_dims(Ainv, 1), A.J)
    grad = np.trace(Bt, axis1=-2, axis2=-1)
    lap = np.trace(Ainv @ A.lap, axis1=-2, axis2=-1) - np.sum(Bt * np.swapaxes(Bt, -1, -2), axis=(1, -2, -1))
    return sign, Lap(ld, grad, lap)
