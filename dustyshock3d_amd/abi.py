"""ctypes mirror of include/dsh.h and the loader of the in-tree libdsh.so.

The library is the product path: nothing here falls back to another implementation.
If libdsh.so is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsh.so")

MAX_DUST = 2
MAX_SPECIES = 1 + MAX_DUST

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_STATE = -4
ERR_GRID = -5
ERR_NONFINITE = -6
ERR_IO = -7
ERR_BUDGET = -8
ERR_NOTFOUND = -9
ERR_UNSUPPORTED = -10

DT_FIXED, DT_CFL_ONCE, DT_CFL = 0, 1, 2
FP64, MIXED = 0, 1
PAIR_DIRECT, PAIR_TILED = 0, 1

_D2 = C.c_double * 2


class Params(C.Structure):
    """dsh_params: Init.txt lines 2-18 (kernel.cu:1705-1722) followed by extensions."""

    _fields_ = [
        ("x_min", C.c_double), ("x_max", C.c_double),
        ("y_min", C.c_double), ("y_max", C.c_double),
        ("z_min", C.c_double), ("z_max", C.c_double),
        ("eos_type", C.c_int), ("max_particles", C.c_int),
        ("particles_per_length", C.c_double), ("h", C.c_double),
        ("avg_cell_coeff", C.c_double), ("tau", C.c_double),
        ("t_end", C.c_double), ("t_out", C.c_double),
        ("alpha", C.c_double), ("beta", C.c_double), ("eps", C.c_double),
        ("periodic_y", C.c_int), ("periodic_z", C.c_int),
        ("gamma", C.c_double), ("cs", C.c_double),
        ("h_mode", C.c_int), ("eta", C.c_double),
        ("n_dust", C.c_int), ("drag_mode", C.c_int),
        ("K", _D2), ("R_dust", _D2), ("mass_ratio", _D2), ("eps_fallback", _D2),
        ("tstop_fallback", C.c_double), ("eps2_uses_count1", C.c_int),
        ("dt_mode", C.c_int), ("cfl", C.c_double), ("max_steps", C.c_int),
        ("cell_size", C.c_double), ("cell_slack", C.c_double),
        ("avg_factor", C.c_int), ("precision", C.c_int),
        ("fp64_wall_band", C.c_int), ("wall_band_speed", C.c_double), ("wall_band_horizon", C.c_int),
    ]


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class SoaView(C.Structure):
    """dsh_soa_view: the data.h structure-of-arrays layout (data.h:19-20,49-55)."""

    _fields_ = [
        ("x", _DP), ("y", _DP), ("z", _DP),
        ("vx", _DP), ("vy", _DP), ("vz", _DP),
        ("ax", _DP), ("ay", _DP), ("az", _DP),
        ("m", _DP), ("rho", _DP), ("p", _DP), ("e", _DP), ("h", _DP),
        ("t_stop", _DP), ("edot", _DP),
        ("ind", _IP), ("nn", _IP), ("cell", _IP),
    ]


class TubeSpec(C.Structure):
    """dsh_tube_spec: one isothermal dusty shock tube."""

    _fields_ = [
        ("x_min", C.c_double), ("x_max", C.c_double),
        ("y_width", C.c_double), ("z_width", C.c_double),
        ("dx_left", C.c_double), ("rho_left", C.c_double),
        ("dust_to_gas", C.c_double), ("K", C.c_double),
        ("jitter", C.c_double), ("frozen_h", C.c_double),
        ("t_end", C.c_double), ("cfl", C.c_double),
        ("dt_mode", C.c_int), ("max_steps", C.c_int), ("precision", C.c_int),
        ("tube_index", C.c_int), ("seed", C.c_uint64),
    ]


CLOCK_FN = C.CFUNCTYPE(C.c_double, C.c_void_p)


class ServerOpts(C.Structure):
    """dsh_server_opts."""

    _fields_ = [
        ("hbm_budget", C.c_double), ("hbm_capacity", C.c_double),
        ("steps_per_epoch", C.c_int), ("backend", C.c_int), ("pair_variant", C.c_int),
        ("clock", CLOCK_FN), ("clock_user", C.c_void_p), ("fake_fail_rebuild", C.c_int),
        ("heavy_work", C.c_double), ("max_heavy", C.c_int),
    ]


BACKEND_GPU, BACKEND_FAKE = 0, 1
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_CANCELLED, REQ_FAILED, REQ_REJECTED = range(6)

DOUBLE_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "rho", "p", "e", "h",
                 "t_stop", "edot")
INT_FIELDS = ("ind", "nn", "cell")
ALL_FIELDS = DOUBLE_FIELDS + INT_FIELDS


def empty_arrays(n: int, fields=ALL_FIELDS) -> dict:
    out = {}
    for f in fields:
        out[f] = np.zeros(n, dtype=np.int32 if f in INT_FIELDS else np.float64)
    return out


def field_arrays(arrays: dict, fields, n: int = None) -> dict:
    """Contiguous int32 / float64 copies (or the arrays themselves) of the given fields that are
    present, each checked to be one-dimensional with n entries (n: the length of x).  The C ABI
    reads n entries of every field it is given, so a shorter array would be read past its end:
    a length mismatch raises ValueError instead."""
    if n is None:
        if "x" not in arrays or arrays["x"] is None:
            raise ValueError("the arrays need an x field")
        n = len(arrays["x"])
    out = {}
    for f in fields:
        a = arrays.get(f)
        if a is None:
            continue
        a = np.ascontiguousarray(a, dtype=np.int32 if f in INT_FIELDS else np.float64)
        if a.ndim != 1 or a.shape[0] != n:
            raise ValueError(f"field {f!r} has shape {a.shape}, expected ({n},)")
        out[f] = a
    return out


def make_view(arrays: dict) -> SoaView:
    """Build a view over numpy arrays (the arrays must stay alive while the view is used)."""
    v = SoaView()
    for f, a in arrays.items():
        if a is None:
            continue
        want = np.int32 if f in INT_FIELDS else np.float64
        if a.dtype != want or not a.flags.c_contiguous:
            raise TypeError(f"field {f!r} must be a contiguous {np.dtype(want).name} array")
        setattr(v, f, a.ctypes.data_as(_IP if f in INT_FIELDS else _DP))
    return v


def spec_dict(spec: TubeSpec) -> dict:
    return {k: getattr(spec, k) for k, _ in TubeSpec._fields_}


def params_dict(p: Params) -> dict:
    out = {}
    for k, _ in Params._fields_:
        v = getattr(p, k)
        out[k] = list(v) if isinstance(v, C.Array) else v
    return out


_LIB = None


def _declare(lib):
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "dsh_default_params": (C.c_int, [P(Params)]),
        "dsh_struct_sizes": (C.c_int, [P(C.c_size_t)]),
        "dsh_tube_params": (C.c_int, [P(TubeSpec), P(Params)]),
        "dsh_tube_count": (C.c_int, [P(TubeSpec), _IP]),
        "dsh_tube_ic": (C.c_int, [P(TubeSpec), C.c_int, P(SoaView)]),
        "dsh_read_init": (C.c_int, [C.c_char_p, P(Params)]),
        "dsh_reference_setup": (C.c_int, [P(Params), P(C.c_int)]),
        "dsh_reference_ic": (C.c_int, [P(Params), C.c_int, P(SoaView)]),
        "dsh_write_data": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                     P(SoaView), C.c_int]),
        "dsh_reference_nn": (C.c_int, [P(Params), P(C.c_double), P(C.c_double), P(C.c_double), C.c_int,
                                       P(C.c_int)]),
        "dsh_create": (C.c_int, [P(Params), P(vp)]),
        "dsh_upload_species": (C.c_int, [vp, C.c_int, P(SoaView), C.c_int]),
        "dsh_step": (C.c_int, [vp, C.c_int, P(C.c_double)]),
        "dsh_download_species": (C.c_int, [vp, C.c_int, P(SoaView)]),
        "dsh_conservation": (C.c_int, [vp, _DP]),
        "dsh_status": (C.c_int, [vp, _DP, _DP, _IP, _IP]),
        "dsh_grid": (C.c_int, [vp, _DP, _IP]),
        "dsh_set_stats": (C.c_int, [vp, C.c_int]),
        "dsh_set_tile": (C.c_int, [vp, C.c_int]),
        "dsh_autotune": (C.c_int, [vp, C.c_int, C.c_int, P(C.c_int), P(C.c_double)]),
        "dsh_batch_set_tile": (C.c_int, [vp, C.c_int]),
        "dsh_set_block_waves": (C.c_int, [vp, C.c_int]),
        "dsh_block_shape": (C.c_int, [vp, P(C.c_int)]),
        "dsh_batch_set_block_waves": (C.c_int, [vp, C.c_int]),
        "dsh_batch_block_shape": (C.c_int, [vp, P(C.c_int)]),
        "dsh_batch_autotune": (C.c_int, [vp, C.c_int, C.c_int, P(C.c_int), P(C.c_double)]),
        "dsh_set_pair_variant": (C.c_int, [vp, C.c_int]),
        "dsh_neighbour_stats": (C.c_int, [vp, C.c_int, _IP, P(C.c_uint64)]),
        "dsh_sorted_ids": (C.c_int, [vp, C.c_int, _IP]),
        "dsh_set_timing": (C.c_int, [vp, C.c_int]),
        "dsh_kernel_times": (C.c_int, [vp, _DP, C.c_char_p, C.c_int, C.c_int]),
        "dsh_last_error": (C.c_char_p, [vp]),
        "dsh_destroy": (None, [vp]),
        "dsh_step_noalloc": (C.c_int, [P(Params), P(SoaView), _IP, C.c_int, P(SoaView), _DP, _DP, vp]),
        "dsh_batch_create": (C.c_int, [P(Params), C.c_int, P(vp)]),
        "dsh_batch_from_specs": (C.c_int, [P(TubeSpec), C.c_int, P(vp)]),
        "dsh_batch_upload": (C.c_int, [vp, C.c_int, C.c_int, P(SoaView), C.c_int]),
        "dsh_batch_run": (C.c_int, [vp, C.c_int, _IP]),
        "dsh_batch_status": (C.c_int, [vp, C.c_int, _DP, _DP, _IP, _IP]),
        "dsh_batch_download": (C.c_int, [vp, C.c_int, C.c_int, P(SoaView)]),
        "dsh_batch_conservation": (C.c_int, [vp, C.c_int, _DP]),
        "dsh_batch_set_pair_variant": (C.c_int, [vp, C.c_int]),
        "dsh_batch_set_timing": (C.c_int, [vp, C.c_int]),
        "dsh_batch_kernel_times": (C.c_int, [vp, _DP, C.c_char_p, C.c_int, C.c_int]),
        "dsh_batch_device_bytes": (C.c_int, [vp, _DP]),
        "dsh_batch_last_error": (C.c_char_p, [vp]),
        "dsh_batch_destroy": (None, [vp]),
        "dsh_server_create": (C.c_int, [P(ServerOpts), P(vp)]),
        "dsh_server_submit": (C.c_int, [vp, P(TubeSpec), C.c_double, P(C.c_uint64)]),
        "dsh_server_pump": (C.c_int, [vp, _IP, _IP]),
        "dsh_server_poll": (C.c_int, [vp, C.c_uint64, _IP, _DP, _DP, _DP, _DP]),
        "dsh_server_result": (C.c_int, [vp, C.c_uint64, C.c_int, P(SoaView)]),
        "dsh_server_cancel": (C.c_int, [vp, C.c_uint64]),
        "dsh_server_footprint": (C.c_int, [P(TubeSpec), _DP]),
        "dsh_server_work": (C.c_int, [P(TubeSpec), _DP]),
        "dsh_server_stats": (C.c_int, [vp, _DP, _IP, _IP]),
        "dsh_server_times": (C.c_int, [vp, _DP, _DP]),
        "dsh_server_last_error": (C.c_char_p, [vp]),
        "dsh_server_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """Load the in-tree libdsh.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _LIB = _declare(C.CDLL(LIB_PATH))
    return _LIB


class DshError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dsh error {code}: {msg}")
        self.code = code
