"""ctypes mirror of include/revsim.h and the loader of the in-tree librevsim.so.

The configuration struct is laid out field for field as ``rs_config`` in ``include/revsim.h``;
``rs_config_size()`` exported by the library is checked against ``ctypes.sizeof`` at load time.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

RS_ABI_VERSION = 2
RS_PRECISION_FP64 = 0
RS_PRECISION_FP32 = 1

ERRORS = {
    -1: "RS_E_CONFIG",
    -2: "RS_E_HIP",
    -3: "RS_E_OVERFLOW",
    -4: "RS_E_NONFINITE",
    -5: "RS_E_STALE",
    -6: "RS_E_STATE",
    -7: "RS_E_NODEVICE",
}

PHASES = ("fwd_pair", "fwd_step", "rev_pair", "rev_step", "list", "loss", "mlp")
MLP_NW = 5313

_i32 = ctypes.c_int32
_f64 = ctypes.c_double


class RsConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", _i32),
        ("n_molecules", _i32),
        ("box", _f64 * 3),
        ("cutoff", _f64),
        ("skin", _f64),
        ("nl_every", _i32),
        ("precision", _i32),
        ("dt", _f64),
        ("gamma", _f64),
        ("temperature", _f64),
        ("eps_rf", _f64),
        ("coulomb_k", _f64),
        ("kB", _f64),
        ("mass_O", _f64),
        ("mass_H", _f64),
        ("theta0", _f64),
        ("rdf_bins", _i32),
        ("rdf_rmax", _f64),
        ("rdf_width", _f64),
        ("rdf_window", _f64),
        ("w_rdf", _f64),
        ("n_steps_loss", _i32),
        ("n_steps_trunc", _i32),
        ("n_steps_log", _i32),
        ("msd_window_steps", _i32),
        ("msd_every", _i32),
        ("msd_target", _f64),
        ("w_msd", _f64),
        ("clip_norm", _f64),
        ("use_mlp", _i32),
        ("pair_chunks", _i32),
        ("w_dhvap", _f64),
        ("dhvap_target", _f64),
        ("dhvap_u_gas", _f64),
        ("barostat_every", _i32),
        ("barostat_pressure", _f64),
        ("reserved", _i32 * 8),
    ]


def make_config(n_molecules: int, box, cutoff: float, **kw) -> RsConfig:
    """Water defaults (water/water.jl:27-49 constants) with overrides."""
    c = RsConfig()
    c.abi_version = RS_ABI_VERSION
    c.n_molecules = int(n_molecules)
    b = [float(box)] * 3 if np.isscalar(box) else [float(v) for v in box]
    for k in range(3):
        c.box[k] = b[k]
    c.cutoff = float(cutoff)
    c.skin = 0.06
    c.nl_every = 10
    c.precision = RS_PRECISION_FP64
    c.dt = 0.001
    c.gamma = 1.0
    c.temperature = 300.0
    c.eps_rf = 78.3
    c.coulomb_k = 138.935457644
    c.kB = 0.00831446262
    c.mass_O = 15.99943
    c.mass_H = 1.007947
    c.theta0 = 1.82421813418
    c.rdf_bins = 200
    c.rdf_rmax = 0.8
    c.rdf_width = 0.0005
    c.rdf_window = 0.04
    c.w_rdf = 1.0
    c.n_steps_loss = 0
    c.n_steps_trunc = 0
    c.n_steps_log = 0
    c.msd_window_steps = 0
    c.msd_every = 0
    c.msd_target = 2.3e-9
    c.w_msd = 0.0
    c.clip_norm = math.inf
    c.use_mlp = 0
    c.pair_chunks = 0
    c.w_dhvap = 0.0
    c.dhvap_target = 44.12 + 2.8  # water/water.jl:751
    c.dhvap_u_gas = 3.69          # water/water.jl:741
    c.barostat_every = 0
    c.barostat_pressure = 1.0     # bar (water/water.jl:1136)
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(f"unknown rs_config field {k!r}")
        setattr(c, k, v)
    return c


def config_dict(c: RsConfig) -> dict:
    out = {}
    for name, _ in RsConfig._fields_:
        v = getattr(c, name)
        if name in ("box", "reserved"):
            v = list(v)
        out[name] = v
    return out


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "librevsim.so"


def library_path() -> str:
    return os.path.join(PKG_DIR, LIB_NAME)


_P = ctypes.POINTER
_dp = _P(ctypes.c_double)
_fp = _P(ctypes.c_float)
_ip = _P(ctypes.c_int32)
_lp = _P(ctypes.c_int64)
_vp = ctypes.c_void_p

# (name, restype, argtypes) for every entry point declared in include/revsim.h
SIGNATURES = [
    ("rs_default_config", None, [_P(RsConfig)]),
    ("rs_create", ctypes.c_int, [_P(RsConfig), ctypes.c_int32, _P(_vp)]),
    ("rs_destroy", ctypes.c_int, [_vp]),
    ("rs_last_error", ctypes.c_char_p, [_vp]),
    ("rs_version", ctypes.c_int, []),
    ("rs_config_size", ctypes.c_int32, []),
    ("rs_set_theta", ctypes.c_int, [_vp, _dp]),
    ("rs_set_target_rdf", ctypes.c_int, [_vp, _dp, ctypes.c_int32]),
    ("rs_set_reverse_options", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double]),
    ("rs_set_external_noise", ctypes.c_int, [_vp, _dp, ctypes.c_int32]),
    ("rs_forward", ctypes.c_int, [_vp, _dp, _dp, ctypes.c_uint64, ctypes.c_int32, _dp, _dp, _dp]),
    ("rs_reverse", ctypes.c_int,
     [_vp, _dp, _dp, ctypes.c_uint64, ctypes.c_int32, _dp, _dp, _dp, _dp]),
    ("rs_value_and_grad", ctypes.c_int,
     [_vp, _dp, _dp, ctypes.c_uint64, ctypes.c_int32, _dp, _dp, _fp]),
    ("rs_set_mlp", ctypes.c_int, [_vp, _fp, ctypes.c_int64]),
    ("rs_mlp_grad", ctypes.c_int, [_vp, _dp, ctypes.c_int64]),
    ("rs_mlp_table", ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int32, _ip, _dp]),
    ("rs_forces", ctypes.c_int, [_vp, _dp, _dp]),
    ("rs_energy", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    ("rs_reweight", ctypes.c_int,
     [_vp, _dp, _dp, ctypes.c_uint64, ctypes.c_int32, _dp, _dp, _dp, _dp]),
    ("rs_adjoint_terms", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("rs_noise", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_int32, _dp]),
    ("rs_rdf", ctypes.c_int, [_vp, _dp, _dp, _lp, _dp, _dp]),
    ("rs_msd", ctypes.c_int, [_vp, _dp, ctypes.c_int32, _dp, _dp, _dp]),
    ("rs_msd_noalloc", ctypes.c_int, [_vp, _vp, ctypes.c_int32, _dp, _dp, _vp]),
    ("rs_molecule_pairs", ctypes.c_int, [_vp, _dp, _ip, ctypes.c_int64, _lp]),
    ("rs_list_radius", ctypes.c_double, [_vp]),
    ("rs_loss_terms", ctypes.c_int, [_vp, _dp]),
    ("rs_stats", ctypes.c_int, [_vp, _lp, ctypes.c_int32]),
    ("rs_get_box", ctypes.c_int, [_vp, _dp]),
    ("rs_set_box", ctypes.c_int, [_vp, _dp]),
    ("rs_timings", ctypes.c_int, [_vp, _dp, ctypes.c_int32]),
    ("rs_timing_counts", ctypes.c_int, [_vp, _lp, ctypes.c_int32]),
    ("rs_enable_timing", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("rs_copy_calibration", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32]),
]


def declared_symbols() -> list[str]:
    """Names of the entry points declared in include/revsim.h (parsed from the header)."""
    import re

    hdr = os.path.join(PKG_DIR, "..", "..", "include", "revsim.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(rs_\w+)\s*\(", txt, flags=re.M)))


_LIB = None


class LibraryMissing(RuntimeError):
    pass


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load librevsim.so (built in-tree by __graft_entry__.build()). Raises if it is absent:
    there is no CPU fallback behind this API."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or library_path()
    if not os.path.exists(p):
        raise LibraryMissing(
            f"{p} not found: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.rs_config_size() != ctypes.sizeof(RsConfig):
        raise RuntimeError(
            f"rs_config layout mismatch: library {lib.rs_config_size()} B vs ctypes {ctypes.sizeof(RsConfig)} B"
        )
    if path is None:
        _LIB = lib
    return lib


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def lptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_lp)
