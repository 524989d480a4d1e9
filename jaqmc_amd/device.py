"""Device-resident wavefunction model over the C ABI.

``DeviceModel`` owns one ``jq_model`` handle: packed weights in HBM plus the
walker-chunk workspace.  PyTorch is used only for device memory and the
current HIP stream; every computation is a HIP kernel of libjaqmc_amd.so.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .params import ModelSpec, pack


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class DeviceModel:
    """A FermiNet or Psiformer wavefunction on one GPU.

    precision: arithmetic of the layer / tangent GEMMs.  "fp32" (FP32 MFMA);
    "bf16" (one BF16 term per operand); "bf16x3" (two BF16 terms, three
    products, ~2^-16); "bf16x6" (three BF16 terms, six products, FP32-level);
    "fp16x3" (two FP16 terms, three products, ~2^-22, power-of-two row and
    column scaling for FP16's range; value-only launches use bf16x6).  Every
    mode accumulates in FP32 and keeps activations FP32 in HBM; determinants,
    E_L assembly and reductions are FP64 in every mode.
    """

    def __init__(self, spec: ModelSpec, atoms, charges, params, precision="fp32", chunk_walkers=0,
                 with_gradients=False):
        lib = _lib.load()
        if not torch.cuda.is_available():
            raise _lib.JqError("DeviceModel needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.spec = spec
        self.N = spec.n_elec
        self.precision = precision
        self.atoms = np.ascontiguousarray(atoms, np.float32).reshape(-1, 3)
        self.charges = np.ascontiguousarray(charges, np.float32).reshape(-1)
        d = _lib.JqModelDesc()
        d.kind = _lib.KIND[spec.kind]
        d.n_up, d.n_dn = spec.nspins
        d.n_atoms = self.atoms.shape[0]
        d.ndets = spec.ndets
        if spec.kind == "ferminet":
            d.n_layers = len(spec.hidden_single)
            for i, (a, b) in enumerate(zip(spec.hidden_single, spec.hidden_double)):
                d.hidden_single[i] = a
                d.hidden_double[i] = b
        else:
            d.n_layers = spec.num_layers
            d.num_heads = spec.num_heads
            d.heads_dim = spec.heads_dim
            d.n_mlp = len(spec.mlp_hidden)
            for i, h in enumerate(spec.mlp_hidden):
                d.mlp_hidden[i] = h
            d.jastrow = 1 if spec.jastrow else 0
        d.envelope = _lib.ENVELOPE[spec.envelope]
        d.precision = _lib.PRECISION[precision]
        d.layer_norm_mode = _lib.LAYER_NORM[spec.layer_norm_mode] if spec.kind == "psiformer" else 0
        d.options = spec.options() | (_lib.OPT_WITH_GRADIENTS if with_gradients else 0)
        d.chunk_walkers = int(chunk_walkers)
        d.atoms = self.atoms.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        d.charges = self.charges.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        flat = pack(spec, params)
        h = ctypes.c_void_p()
        torch.cuda.current_device()  # make sure the HIP context exists on the current device
        _lib.check(lib.jq_model_create(ctypes.byref(d), flat.ctypes.data_as(ctypes.c_void_p), flat.size,
                                       ctypes.byref(h)), "jq_model_create")
        self._h = h
        self._lib = lib
        self.device = torch.device("cuda", torch.cuda.current_device())

    @property
    def workspace_bytes(self):
        return int(self._lib.jq_model_workspace_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.jq_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_elec(self, elec):
        if not (elec.is_cuda and elec.dtype == torch.float32 and elec.is_contiguous()):
            raise ValueError("electrons must be a contiguous float32 CUDA tensor")
        if elec.ndim != 3 or elec.shape[1] != self.N or elec.shape[2] != 3:
            raise ValueError(f"electrons must have shape (B, {self.N}, 3), got {tuple(elec.shape)}")
        return elec.shape[0]

    # -- entry points ----------------------------------------------------------
    def logpsi(self, elec):
        B = self._check_elec(elec)
        lp = torch.empty(B, dtype=torch.float32, device=elec.device)
        sg = torch.empty(B, dtype=torch.float32, device=elec.device)
        _lib.check(self._lib.jq_logpsi(self._h, _ptr(elec), B, _ptr(lp), _ptr(sg), _stream()), "jq_logpsi")
        return lp, sg

    def local_energy(self, elec, with_grad=False):
        B = self._check_elec(elec)
        dev = elec.device
        out = {k: torch.empty(B, dtype=torch.float64, device=dev) for k in ("e_kin", "e_pot", "e_loc")}
        out["logpsi"] = torch.empty(B, dtype=torch.float32, device=dev)
        grad = torch.empty(B, 3 * self.N, dtype=torch.float32, device=dev) if with_grad else None
        _lib.check(self._lib.jq_local_energy(self._h, _ptr(elec), B, _ptr(out["e_kin"]), _ptr(out["e_pot"]),
                                             _ptr(out["e_loc"]), _ptr(out["logpsi"]), _ptr(grad), _stream()),
                   "jq_local_energy")
        if with_grad:
            out["grad"] = grad
        return out

    @property
    def n_params(self):
        return int(self._lib.jq_model_param_count(self._h))

    def logpsi_vjp(self, elec, cot):
        """sum_w cot[w] d log|psi(x_w)| / d theta as a float32 vector in pack
        order (models created with with_gradients=True; Psiformer with pre-LN
        or no LayerNorm)."""
        B = self._check_elec(elec)
        if not (cot.is_cuda and cot.dtype == torch.float64 and cot.numel() == B and cot.is_contiguous()):
            raise ValueError("cot must be a contiguous float64 CUDA tensor of B entries")
        g = torch.empty(self.n_params, dtype=torch.float32, device=elec.device)
        _lib.check(self._lib.jq_logpsi_vjp(self._h, _ptr(elec), B, _ptr(cot), _ptr(g), _stream()), "jq_logpsi_vjp")
        return g

    def orbitals(self, elec):
        """(B, K, N, N) orbital matrices [walker, det, electron, orbital] (FP32)."""
        B = self._check_elec(elec)
        K = self.spec.ndets
        out = torch.empty(B, K, self.N, self.N, dtype=torch.float32, device=elec.device)
        _lib.check(self._lib.jq_orbitals(self._h, _ptr(elec), B, _ptr(out), _stream()), "jq_orbitals")
        return out

    def sample(self, elec, state, n_outer, steps, adapt_frequency=100, pmove_range=(0.5, 0.55), seed=0, ctr=0,
               n_accept=None, logp=None, pmove_out=None):
        """n_outer device-side MCMCSampler.step calls (moves + width adaptation,
        no host synchronisation).  ``state`` is the float64 CUDA tensor
        {stddev, counter, accepted total, pmoves[adapt_frequency]}; returns
        (logp, n_accept)."""
        B = self._check_elec(elec)
        if not (state.is_cuda and state.dtype == torch.float64 and state.numel() == 3 + adapt_frequency):
            raise ValueError("state must be a float64 CUDA tensor of 3 + adapt_frequency entries")
        if n_accept is None:
            n_accept = torch.zeros(B, dtype=torch.int32, device=elec.device)
        if logp is None:
            logp = torch.empty(B, dtype=torch.float32, device=elec.device)
        if pmove_out is not None and not (pmove_out.dtype == torch.float64 and pmove_out.numel() >= n_outer):
            raise ValueError("pmove_out must be a float64 CUDA tensor of n_outer entries")
        _lib.check(self._lib.jq_sample(self._h, _ptr(elec), _ptr(logp), B, int(n_outer), int(steps), _ptr(state),
                                       int(adapt_frequency), float(pmove_range[0]), float(pmove_range[1]),
                                       ctypes.c_uint64(seed), ctypes.c_uint64(ctr), _ptr(n_accept), _ptr(pmove_out),
                                       _stream()), "jq_sample")
        return logp, n_accept

    def mh_steps(self, elec, logp, n_steps, width, normals=None, uniforms=None, seed=0, ctr=0, n_accept=None):
        """In-place Metropolis moves; returns the int32 accept counts (B,)."""
        B = self._check_elec(elec)
        if n_accept is None:
            n_accept = torch.zeros(B, dtype=torch.int32, device=elec.device)
        for t in (normals, uniforms):
            if t is not None and not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                raise ValueError("normals / uniforms must be contiguous float32 CUDA tensors")
        if normals is not None and normals.numel() != n_steps * B * 3 * self.N:
            raise ValueError("normals must have shape (n_steps, B, N, 3)")
        if uniforms is not None and uniforms.numel() != n_steps * B:
            raise ValueError("uniforms must have shape (n_steps, B)")
        _lib.check(self._lib.jq_mh_steps(self._h, _ptr(elec), _ptr(logp), B, int(n_steps), float(width),
                                         _ptr(normals), _ptr(uniforms), ctypes.c_uint64(seed), ctypes.c_uint64(ctr),
                                         _ptr(n_accept), _stream()), "jq_mh_steps")
        return n_accept


def reduce_mean_var(x):
    """(mean, var, count) of the finite entries of a float64 CUDA tensor, in FP64 on the GPU."""
    lib = _lib.load()
    out = torch.empty(3, dtype=torch.float64, device=x.device)
    _lib.check(lib.jq_reduce_mean_var(_ptr(x), x.numel(), _ptr(out), _stream()), "jq_reduce_mean_var")
    return out


CLIP_METHODS = {"mad": 0, "iqr": 1, "none": 2}


def chunk_for_budget(spec: ModelSpec, atoms, charges, params, precision, walkers, budget_bytes):
    """Walkers per workspace chunk for a batch of `walkers` whose model should
    hold at most `budget_bytes` of device memory: the workspace is affine in
    the chunk (probed with chunks of one and two walkers), and the batch is
    split into equal chunks of at most the largest chunk that fits."""
    ws = []
    for c in (1, 2):
        m = DeviceModel(spec, atoms, charges, params, precision=precision, chunk_walkers=c)
        ws.append(m.workspace_bytes)
        m.close()
    return equal_chunks(ws[0], ws[1], walkers, budget_bytes)


def equal_chunks(ws1: int, ws2: int, walkers: int, budget_bytes) -> int:
    """Walkers per chunk from the workspace bytes at chunks of one and two
    walkers (affine in the chunk): the fewest equal chunks whose workspace fits
    budget_bytes (at least one walker per chunk)."""
    per = max(1, int(ws2) - int(ws1))
    fixed = int(ws1) - per
    cmax = max(1, (int(budget_bytes) - fixed) // per)
    n = -(-int(walkers) // cmax)
    return -(-int(walkers) // n)


def loss_clip(x, method="mad", scale=5.0, workspace=None):
    """LossAndGrad's clipping on the GPU (FP64): returns (clipped, weights,
    stats) with weights = 2 (clipped - mean(clipped)) / n and stats =
    {nanmean(x), nanmean(clipped), lo, hi, n}."""
    lib = _lib.load()
    B = x.numel()
    nb = int(lib.jq_loss_clip_workspace_bytes(B))
    if workspace is None or workspace.numel() < nb:
        workspace = torch.empty(nb, dtype=torch.uint8, device=x.device)
    clipped = torch.empty(B, dtype=torch.float64, device=x.device)
    w = torch.empty(B, dtype=torch.float64, device=x.device)
    stats = torch.empty(5, dtype=torch.float64, device=x.device)
    _lib.check(lib.jq_loss_clip(_ptr(x), B, CLIP_METHODS[method], float(scale), _ptr(clipped), _ptr(w), _ptr(stats),
                                _ptr(workspace), nb, _stream()), "jq_loss_clip")
    return clipped, w, stats


def loss_clip_noalloc(x, method="mad", scale=5.0, out=None):
    """loss_clip through jq_loss_clip_noalloc: the library allocates nothing
    (its scratch is a static device array per GPU); out = (clipped, weights,
    stats) preallocated tensors, or None to allocate them here."""
    lib = _lib.load()
    B = x.numel()
    if out is None:
        out = (torch.empty(B, dtype=torch.float64, device=x.device),
               torch.empty(B, dtype=torch.float64, device=x.device),
               torch.empty(5, dtype=torch.float64, device=x.device))
    clipped, w, stats = out
    _lib.check(lib.jq_loss_clip_noalloc(_ptr(x), B, CLIP_METHODS[method], float(scale), _ptr(clipped), _ptr(w),
                                        _ptr(stats), _stream()), "jq_loss_clip_noalloc")
    return clipped, w, stats


def reduce_accepts(n_accept):
    lib = _lib.load()
    out = torch.empty(1, dtype=torch.float64, device=n_accept.device)
    _lib.check(lib.jq_reduce_accepts(_ptr(n_accept), n_accept.numel(), _ptr(out), _stream()), "jq_reduce_accepts")
    return out
