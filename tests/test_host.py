"""Host logic and the C-ABI library's exports (CPU only)."""

import ctypes
import re

import numpy as np
import pytest

from jaqmc_amd import _lib
from jaqmc_amd import params as P
from jaqmc_amd import systems as S


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    import pathlib

    root = pathlib.Path(__file__).resolve().parent.parent
    names = set()
    for h in (root / "include").glob("*.h"):
        names |= set(re.findall(r"\b(jq_[a-z_]+)\s*\(", h.read_text()))
    assert "jq_local_energy" in names and "jq_mh_steps" in names
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_typed_binding_loads():
    lib = _lib.load()
    assert lib.jq_last_error() is not None


@pytest.mark.parametrize("name", ["li", "n2", "ring42", "h60"])
def test_param_pack_roundtrip(name):
    sysm, spec, *_ = S.workload(name)
    p = P.init_params(spec, seed=3, bias_scale=0.1)
    flat = P.pack(spec, p)
    q = P.unpack(spec, flat, p)
    np.testing.assert_array_equal(P.pack(spec, q), flat)
    assert flat.dtype == np.float32


def test_ferminet_layer_widths():
    # FermiLayers.aggregate_features: in_l = f1 (1 + nch) + f2 nch
    _, spec, *_ = S.workload("n2")
    ins, ins2 = spec.ferminet_dims()
    assert ins == [4 * 2 * 3 + 4 * 2, 256 * 3 + 32 * 2, 832, 832]
    assert ins2 == [4, 32, 32]
    _, li, *_ = S.workload("li")
    assert li.ferminet_dims()[0] == [4 * 3 + 8, 64 * 3 + 16 * 2]


def test_walker_init_order_and_geometry():
    s = S.n2_molecule()
    assert s.spins_per_atom == [(4, 3), (3, 4)] or sum(a for a, _ in s.spins_per_atom) == 7
    x = S.init_walkers(s, 64, seed=0, width=1e-3)
    # electrons ordered all-up (atom 0, atom 1) then all-down
    ups = [0] * s.spins_per_atom[0][0] + [1] * s.spins_per_atom[1][0]
    for i, a in enumerate(ups):
        np.testing.assert_allclose(x[:, i].mean(0), s.atoms[a], atol=1e-2)
    r = S.ring42()
    assert r.atoms.shape == (12, 3) and r.nspins == (21, 21)
    np.testing.assert_allclose(np.linalg.norm(r.atoms[:6], axis=1), 2.64, rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(r.atoms[6:], axis=1), 4.68, rtol=1e-6)
    h = S.h_chain()
    assert h.nspins == (30, 30) and h.atoms[1, 0] == pytest.approx(1.8)


def test_random_chain_is_neutral_and_spin_balanced():
    rng = np.random.default_rng(0)
    for n in range(2, 43):
        s = S.random_chain(rng, n)
        assert int(s.charges.sum()) == n
        assert s.nspins == ((n + 1) // 2, n // 2)
        assert s.charges.max() <= 8 and s.charges.min() >= 1
        if len(s.atoms) > 1:
            d = np.linalg.norm(np.diff(s.atoms, axis=0), axis=1)
            assert d.min() >= 1.5 - 1e-6 and d.max() <= 3.0 + 1e-6


def test_device_model_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from jaqmc_amd.device import DeviceModel

    sysm, spec, *_ = S.workload("li")
    with pytest.raises(_lib.JqError):
        DeviceModel(spec, sysm.atoms, sysm.charges, P.init_params(spec))


def _desc(spec, atoms, charges, precision="fp32", ln="pre"):
    d = _lib.JqModelDesc()
    d.kind = _lib.KIND[spec.kind]
    d.n_up, d.n_dn = spec.nspins
    d.n_atoms = atoms.shape[0]
    d.ndets = spec.ndets
    d.n_layers = len(spec.hidden_single) if spec.kind == "ferminet" else spec.num_layers
    for i, (a, b) in enumerate(zip(spec.hidden_single, spec.hidden_double)):
        d.hidden_single[i], d.hidden_double[i] = a, b
    d.num_heads, d.heads_dim, d.n_mlp = spec.num_heads, spec.heads_dim, len(spec.mlp_hidden)
    d.envelope = _lib.ENVELOPE[spec.envelope]
    d.precision = _lib.PRECISION[precision]
    d.layer_norm_mode = _lib.LAYER_NORM[ln]
    d.options = spec.options()
    d.atoms = atoms.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    d.charges = charges.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    return d


@pytest.mark.parametrize("change,msg", [
    (dict(use_last_layer=True), b"use_last_layer"),
    (dict(envelope="diagonal"), b"diagonal"),
    (dict(bias_orbitals=True), b"Psiformer only"),
])
def test_unsupported_options_rejected_before_allocation(change, msg):
    """jq_model_create returns JQ_EINVAL for reference options the device path
    does not implement, before touching the device (runs without a GPU)."""
    lib = _lib.load()
    sysm, spec, *_ = S.workload("li")
    for k, v in change.items():
        setattr(spec, k, v)
    if "bias_orbitals" in change:  # a FermiNet with an orbital bias: the reference never builds one
        spec.kind = "ferminet"
    atoms = np.ascontiguousarray(sysm.atoms, np.float32)
    charges = np.ascontiguousarray(sysm.charges, np.float32)
    d = _desc(spec, atoms, charges)
    if "bias_orbitals" in change:
        d.options = _lib.OPT_ORBITAL_BIAS
    flat = np.zeros(16, np.float32)
    h = ctypes.c_void_p()
    rc = lib.jq_model_create(ctypes.byref(d), flat.ctypes.data_as(ctypes.c_void_p), flat.size, ctypes.byref(h))
    assert rc == -1
    assert msg in lib.jq_last_error()
    assert not h.value


def test_bench_gemm_broadcast_rows_stay_in_bounds():
    """jq_bench_gemm sizes its broadcast term for units // 7 + 2 walkers; the
    unit -> walker map of its 7 + 7 electron layout (csrc/common.h
    walker_of_unit) never reaches past that for any unit count."""
    def walker_of_unit(u, Bc, n_up, n_dn):
        uu = Bc * n_up
        return u // n_up if u < uu else (u - uu) // n_dn

    for units in range(1, 600):
        Bc = units // 14 if units // 14 > 0 else 1
        top = max(walker_of_unit(u, Bc, 7, 7) for u in range(units))
        assert top < units // 7 + 2


def test_equal_chunks_from_workspace_probe():
    """Chunk sizing for a workspace budget: affine workspace (fixed + per
    walker), fewest equal chunks that fit, at least one walker per chunk."""
    from jaqmc_amd.device import equal_chunks

    fixed, per = 1000, 100
    ws1, ws2 = fixed + per, fixed + 2 * per
    assert equal_chunks(ws1, ws2, 2048, fixed + 2048 * per) == 2048  # everything fits
    assert equal_chunks(ws1, ws2, 2048, fixed + 2047 * per) == 1024  # two equal chunks
    assert equal_chunks(ws1, ws2, 2048, fixed + 700 * per) == 683    # three chunks: 683, 683, 682
    assert equal_chunks(ws1, ws2, 10, 0) == 1                         # budget below one walker
    c = equal_chunks(ws1, ws2, 1000, fixed + 333 * per)
    assert fixed + c * per <= fixed + 333 * per and -(-1000 // c) == 4


def test_traffic_summary_calibration_and_exclusions(tmp_path):
    """jaqmc_amd.profiling.traffic: counters scaled by the copy calibration,
    excluded dispatches (by grid size or ordinal) listed but not averaged."""
    import sqlite3

    from jaqmc_amd.profiling import traffic

    def db(path, counter, rows):
        con = sqlite3.connect(path)
        con.execute("create table counters_collection (dispatch_id int, kernel_name text, grid_size int, "
                    "counter_name text, value real)")
        con.executemany("insert into counters_collection values (?, ?, ?, ?, ?)",
                        [(i, n, g, counter, v) for i, (n, g, v) in enumerate(rows)])
        con.commit()
        con.close()

    copy = 1 << 30
    k = "void gemm_kernel<8, 1, 4, 4, true, false>(GemmArgs)"
    # the copy reads as half its bytes on FETCH_SIZE (KiB), in full on WRITE_SIZE
    db(tmp_path / "f.db", "FETCH_SIZE", [("__amd_rocclr_copyBuffer", 1, copy / 2048), (k, 100, 1000.0),
                                         (k, 100, 3000.0), (k, 50, 99999.0)])
    db(tmp_path / "w.db", "WRITE_SIZE", [("__amd_rocclr_copyBuffer", 1, copy / 1024), (k, 100, 10.0),
                                         (k, 100, 30.0), (k, 50, 77777.0)])
    t = traffic(str(tmp_path / "f.db"), str(tmp_path / "w.db"), "gemm_kernel<8, 1, 4, 4, true", copy, ("g50",))
    assert t["FETCH_SIZE"]["scale"] == 2.0 and t["WRITE_SIZE"]["scale"] == 1.0
    assert t["FETCH_SIZE"]["bytes_per_dispatch"] == 2000.0 * 1024 * 2
    assert t["WRITE_SIZE"]["bytes_per_dispatch"] == 20.0 * 1024
    assert len(t["FETCH_SIZE"]["each"]) == 3  # the excluded dispatch is still listed
    t2 = traffic(str(tmp_path / "f.db"), str(tmp_path / "w.db"), "gemm_kernel<8, 1, 4, 4, true", copy, ("d0", "d2"))
    assert t2["FETCH_SIZE"]["bytes_per_dispatch"] == 3000.0 * 1024 * 2
