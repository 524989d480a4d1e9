"""ctypes binding of libjaqmc_amd.so (the C ABI declared in include/jaqmc_amd.h).

The library is loaded from this package directory.  There is no fallback: if
the shared object is missing or a symbol is absent, loading raises, so the
product path can never silently run anywhere but the HIP kernels.
"""

from __future__ import annotations

import ctypes
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libjaqmc_amd.so"
MAX_LAYERS = 8

KIND = {"ferminet": 0, "psiformer": 1}
ENVELOPE = {"isotropic": 0, "abs_isotropic": 1, "null": 2, "diagonal": 3}
PRECISION = {"fp32": 0, "bf16": 1, "bf16x3": 2, "bf16x6": 3, "fp16x3": 4}
LAYER_NORM = {"pre": 0, "post": 1, "null": 2}
# jq_model_desc.options bits (include/jaqmc_amd.h)
OPT_NO_ATTN_BIAS = 1 << 0
OPT_NO_INPUT_BIAS = 1 << 1
OPT_ORBITAL_BIAS = 1 << 2
OPT_NO_RESCALE = 1 << 3
OPT_SHARED_ORBITALS = 1 << 4
OPT_USE_LAST_LAYER = 1 << 5
OPT_WITH_GRADIENTS = 1 << 16

EXPORTS = (
    "jq_model_create", "jq_model_destroy", "jq_model_workspace_bytes", "jq_logpsi", "jq_mh_steps",
    "jq_local_energy", "jq_reduce_mean_var", "jq_reduce_accepts", "jq_last_error",
    "jq_profile_enable", "jq_profile_read", "jq_orbitals", "jq_sample", "jq_loss_clip",
    "jq_loss_clip_workspace_bytes", "jq_logpsi_vjp", "jq_model_param_count", "jq_loss_clip_noalloc",
)

# launch classes of include/jaqmc_amd_prof.h
PROF = {"gemm_layer_lap": 0, "det_lap": 1, "gemm_orb_lap": 2, "pair_lap": 3, "forward": 4,
        "local_energy": 5, "attn_lap": 6}


class JqModelDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("n_up", ctypes.c_int),
        ("n_dn", ctypes.c_int),
        ("n_atoms", ctypes.c_int),
        ("ndets", ctypes.c_int),
        ("n_layers", ctypes.c_int),
        ("hidden_single", ctypes.c_int * MAX_LAYERS),
        ("hidden_double", ctypes.c_int * MAX_LAYERS),
        ("num_heads", ctypes.c_int),
        ("heads_dim", ctypes.c_int),
        ("n_mlp", ctypes.c_int),
        ("mlp_hidden", ctypes.c_int * 4),
        ("jastrow", ctypes.c_int),
        ("envelope", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("chunk_walkers", ctypes.c_int),
        ("atoms", ctypes.POINTER(ctypes.c_float)),
        ("charges", ctypes.POINTER(ctypes.c_float)),
        ("layer_norm_mode", ctypes.c_int),
        ("options", ctypes.c_int),
    ]


class JqError(RuntimeError):
    pass


_lib = None


def load():
    """Load (once) and type the library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise JqError(f"{LIB_PATH} is not built: run `python -m jaqmc_amd.build` (no CPU fallback exists)")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, i32, u64, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_size_t
    lib.jq_model_create.argtypes = [ctypes.POINTER(JqModelDesc), vp, sz, ctypes.POINTER(vp)]
    lib.jq_model_destroy.argtypes = [vp]
    lib.jq_model_workspace_bytes.argtypes = [vp]
    lib.jq_model_workspace_bytes.restype = sz
    lib.jq_logpsi.argtypes = [vp, vp, i32, vp, vp, vp]
    lib.jq_mh_steps.argtypes = [vp, vp, vp, i32, i32, ctypes.c_float, vp, vp, u64, u64, vp, vp]
    lib.jq_local_energy.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp, vp]
    lib.jq_reduce_mean_var.argtypes = [vp, i32, vp, vp]
    lib.jq_reduce_accepts.argtypes = [vp, i32, vp, vp]
    lib.jq_orbitals.argtypes = [vp, vp, i32, vp, vp]
    f64 = ctypes.c_double
    lib.jq_sample.argtypes = [vp, vp, vp, i32, i32, i32, vp, i32, f64, f64, u64, u64, vp, vp, vp]
    lib.jq_loss_clip_workspace_bytes.argtypes = [i32]
    lib.jq_loss_clip_workspace_bytes.restype = sz
    lib.jq_loss_clip.argtypes = [vp, i32, i32, f64, vp, vp, vp, vp, sz, vp]
    lib.jq_loss_clip_noalloc.argtypes = [vp, i32, i32, f64, vp, vp, vp, vp]
    lib.jq_logpsi_vjp.argtypes = [vp, vp, i32, vp, vp, vp]
    lib.jq_model_param_count.argtypes = [vp]
    lib.jq_model_param_count.restype = sz
    lib.jq_last_error.restype = ctypes.c_char_p
    lib.jq_profile_enable.argtypes = [i32]
    lib.jq_profile_read.argtypes = [i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    for name in EXPORTS:
        getattr(lib, name).restype = getattr(lib, name).restype if name in (
            "jq_model_workspace_bytes", "jq_last_error", "jq_loss_clip_workspace_bytes",
            "jq_model_param_count") else ctypes.c_int
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().jq_last_error().decode(errors="replace")
        raise JqError(f"{what} failed ({rc}): {msg}")


def profile_enable(on: bool):
    load().jq_profile_enable(1 if on else 0)


def profile_read(cls: str):
    """(total device ms, launch count) of one instrumented launch class."""
    t = ctypes.c_double()
    n = ctypes.c_int()
    check(load().jq_profile_read(PROF[cls], ctypes.byref(t), ctypes.byref(n)), "jq_profile_read")
    return t.value, n.value
