"""Python host mirror of the reference's per-step protocol over the C ABI.

The reference drives its step from main() over global data.h arrays
(kernel.cu:1793-2566).  ``Simulation`` exposes the same operations: create the
device state, upload each species in the data.h layout, advance the time loop,
download the arrays back in original particle order, read the conservation sums.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import DshError, Params, TubeSpec, empty_arrays, make_view


def _check(rc: int, msg: str = ""):
    if rc != abi.OK:
        raise DshError(rc, msg)


def default_params() -> Params:
    p = Params()
    _check(abi.lib().dsh_default_params(C.byref(p)), "default_params")
    return p


def tube_params(spec: TubeSpec) -> Params:
    p = Params()
    _check(abi.lib().dsh_tube_params(C.byref(spec), C.byref(p)), "tube_params")
    return p


def tube_count(spec: TubeSpec) -> int:
    n = (C.c_int * 3)()
    _check(abi.lib().dsh_tube_count(C.byref(spec), n), "tube_count")
    return int(n[0])


def tube_ic(spec: TubeSpec, species: int) -> dict:
    """Initial data.h arrays of one species of a shock tube (original particle order)."""
    n = tube_count(spec)
    arrs = empty_arrays(n)
    v = make_view(arrs)
    _check(abi.lib().dsh_tube_ic(C.byref(spec), species, C.byref(v)), "tube_ic")
    return arrs


def read_init(path: str) -> Params:
    p = Params()
    _check(abi.lib().dsh_read_init(path.encode(), C.byref(p)), f"read_init {path}")
    return p


def reference_setup(p: Params) -> int:
    """Apply init_Sod_X_Dust's parameter overrides to p; particles per species."""
    n = (C.c_int * 3)()
    _check(abi.lib().dsh_reference_setup(C.byref(p), n), "reference_setup")
    return int(n[0])


def reference_ic(p: Params, species: int) -> dict:
    """init_Sod_X_Dust placement of one species (kernel.cu:1391-1575), index order."""
    q = Params.from_buffer_copy(p)
    n = reference_setup(q)
    arrs = empty_arrays(n)
    v = make_view(arrs)
    _check(abi.lib().dsh_reference_ic(C.byref(p), species, C.byref(v)), "reference_ic")
    return arrs


def write_data(directory: str, num: int, t: float, tau: float, h: float, species: int, arrays: dict):
    """Data_out / Data_out_dust files of one species (kernel.cu:232-826)."""
    n = len(arrays["x"])
    v = make_view(abi.field_arrays(arrays, abi.ALL_FIELDS, n))
    _check(abi.lib().dsh_write_data(directory.encode(), num, t, tau, h, species, C.byref(v), n - 1),
           "write_data")


def reference_nn(p: Params, x, y, z) -> np.ndarray:
    """Next-particle pointers of the reference's lists on its 2h grid."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    if not (x.ndim == y.ndim == z.ndim == 1 and len(x) == len(y) == len(z)):
        raise ValueError("x, y and z must be one-dimensional and of equal length")
    nn = np.empty(len(x), dtype=np.int32)
    dp = C.POINTER(C.c_double)
    _check(abi.lib().dsh_reference_nn(C.byref(p), x.ctypes.data_as(dp), y.ctypes.data_as(dp),
                                      z.ctypes.data_as(dp), len(x) - 1, nn.ctypes.data_as(C.POINTER(C.c_int))),
           "reference_nn")
    return nn


UPLOAD_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "rho", "p", "e", "h", "t_stop", "ind")


class Simulation:
    """One tube on the GPU (a batch of one in the engine)."""

    def __init__(self, params: Params):
        self._lib = abi.lib()
        self.params = params
        self._ctx = C.c_void_p()
        _check(self._lib.dsh_create(C.byref(params), C.byref(self._ctx)), "dsh_create")
        self.counts = {}

    @classmethod
    def from_spec(cls, spec: TubeSpec) -> "Simulation":
        sim = cls(tube_params(spec))
        for s in (0, 1):
            sim.upload(s, tube_ic(spec, s))
        return sim

    def _err(self) -> str:
        m = self._lib.dsh_last_error(self._ctx)
        return m.decode() if m else ""

    def _call(self, rc: int):
        if rc != abi.OK:
            raise DshError(rc, self._err())

    def upload(self, species: int, arrays: dict):
        n = len(arrays["x"])
        keep = abi.field_arrays(arrays, UPLOAD_FIELDS, n)
        v = make_view(keep)
        self._call(self._lib.dsh_upload_species(self._ctx, species, C.byref(v), n - 1))
        self.counts[species] = n

    def step(self, nsteps: int = 1) -> float:
        t = C.c_double(0.0)
        self._call(self._lib.dsh_step(self._ctx, nsteps, C.byref(t)))
        return t.value

    def download(self, species: int, fields=abi.ALL_FIELDS) -> dict:
        arrs = empty_arrays(self.counts[species], fields)
        v = make_view(arrs)
        self._call(self._lib.dsh_download_species(self._ctx, species, C.byref(v)))
        return arrs

    def conservation(self) -> np.ndarray:
        out = np.zeros(4)
        self._call(self._lib.dsh_conservation(self._ctx, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def status(self) -> dict:
        t, dt = C.c_double(), C.c_double()
        steps, done = C.c_int(), C.c_int()
        self._call(self._lib.dsh_status(self._ctx, C.byref(t), C.byref(dt), C.byref(steps), C.byref(done)))
        return {"t": t.value, "dt": dt.value, "steps": steps.value, "done": bool(done.value)}

    def grid(self) -> dict:
        d = np.zeros(9)
        i = np.zeros(5, np.int32)
        self._call(self._lib.dsh_grid(self._ctx, d.ctypes.data_as(C.POINTER(C.c_double)),
                                      i.ctypes.data_as(C.POINTER(C.c_int))))
        return {"origin": d[0:3], "cell": d[3:6], "length": d[6:9], "n": i[0:3], "f": int(i[3]),
                "n_avg": int(i[4])}

    def set_tile(self, tile: int):
        """Candidates staged per LDS piece by the tiled pair kernels (128, 160, 192, 224, 256)."""
        self._call(self._lib.dsh_set_tile(self._ctx, int(tile)))

    def set_block_waves(self, waves: int):
        """Waves (64-target chunks) per block of the tiled pair kernels (1, 2, 4)."""
        self._call(self._lib.dsh_set_block_waves(self._ctx, int(waves)))

    def block_shape(self) -> dict:
        shape = (C.c_int * 2)()
        self._call(self._lib.dsh_block_shape(self._ctx, shape))
        return {"tile": int(shape[0]), "waves": int(shape[1])}

    def autotune(self, steps_per_trial: int = 2, with_variants: bool = False) -> dict:
        """Time each pair-kernel configuration (piece size x waves per block) on the live
        state and keep the fastest."""
        chosen = (C.c_int * 2)()
        ms = C.c_double(0.0)
        self._call(self._lib.dsh_autotune(self._ctx, int(steps_per_trial), int(with_variants), chosen, C.byref(ms)))
        return {"tile": int(chosen[0]), "waves": self.block_shape()["waves"], "variant": int(chosen[1]),
                "ms_per_step": ms.value}

    def set_pair_variant(self, variant: int):
        self._call(self._lib.dsh_set_pair_variant(self._ctx, int(variant)))

    def set_stats(self, on: bool = True):
        self._call(self._lib.dsh_set_stats(self._ctx, int(on)))

    def neighbour_stats(self, species: int):
        n = self.counts[species]
        cnt = np.zeros(n, np.int32)
        h = np.zeros(n, np.uint64)
        self._call(self._lib.dsh_neighbour_stats(self._ctx, species, cnt.ctypes.data_as(C.POINTER(C.c_int)),
                                                 h.ctypes.data_as(C.POINTER(C.c_uint64))))
        return cnt, h

    def sorted_ids(self, species: int) -> np.ndarray:
        """Particle indices in the device's sorted order of the last step."""
        out = np.zeros(self.counts[species], np.int32)
        self._call(self._lib.dsh_sorted_ids(self._ctx, species, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def set_timing(self, on: bool = True):
        self._call(self._lib.dsh_set_timing(self._ctx, int(on)))

    def kernel_times(self) -> dict:
        nmax, ln = 32, 32
        ms = np.zeros(nmax)
        names = C.create_string_buffer(nmax * ln)
        n = self._lib.dsh_kernel_times(self._ctx, ms.ctypes.data_as(C.POINTER(C.c_double)), names, nmax, ln)
        if n < 0:
            raise DshError(n, self._err())
        raw = names.raw
        return {raw[i * ln:(i + 1) * ln].split(b"\0")[0].decode(): float(ms[i]) for i in range(min(n, nmax))}

    def close(self):
        if self._ctx:
            self._lib.dsh_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
