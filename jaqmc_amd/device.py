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

    precision: "fp32" (FP32 MFMA layers), "bf16" (BF16 MFMA operands and
    storage, FP32 accumulation) or "bf16x3" (split-BF16 MFMA, FP32 storage).
    Determinants, E_L assembly and reductions are FP64 in every mode.
    """

    def __init__(self, spec: ModelSpec, atoms, charges, params, precision="fp32", chunk_walkers=0):
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


def reduce_accepts(n_accept):
    lib = _lib.load()
    out = torch.empty(1, dtype=torch.float64, device=n_accept.device)
    _lib.check(lib.jq_reduce_accepts(_ptr(n_accept), n_accept.numel(), _ptr(out), _stream()), "jq_reduce_accepts")
    return out
