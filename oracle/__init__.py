"""ctypes wrapper of liboracle.so, the FP64 OpenMP restatement of the reference step.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker / CPU baseline.  The product (dustyshock3d_amd) never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from dustyshock3d_amd.abi import (ALL_FIELDS, OK, DshError, Params, SoaView, TubeSpec, empty_arrays, field_arrays,
                                  make_view)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        vp = C.c_void_p
        DP, IP = P(C.c_double), P(C.c_int)
        sig = {
            "orc_create": (C.c_int, [P(Params), P(vp)]),
            "orc_upload": (C.c_int, [vp, C.c_int, P(SoaView), C.c_int]),
            "orc_step": (C.c_int, [vp, C.c_int, DP]),
            "orc_download": (C.c_int, [vp, C.c_int, P(SoaView)]),
            "orc_conservation": (C.c_int, [vp, DP]),
            "orc_status": (C.c_int, [vp, DP, DP, IP, IP]),
            "orc_neighbour_stats": (C.c_int, [vp, C.c_int, IP, P(C.c_uint64)]),
            "orc_sorted_ids": (C.c_int, [vp, C.c_int, IP]),
            "orc_grid": (C.c_int, [vp, DP, IP]),
            "orc_num_avg_cells": (C.c_int, [vp]),
            "orc_avg_cells": (C.c_int, [vp, IP, IP, DP, DP, DP, DP, DP]),
            "orc_destroy": (None, [vp]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_get_threads": (C.c_int, []),
            "orc_W": (C.c_double, [C.c_double, C.c_double]),
            "orc_dW": (C.c_double, [C.c_double, C.c_double]),
            "orc_idic_cell": (None, [C.c_int, DP, DP, DP, DP, DP, C.c_double, DP, DP]),
            "orc_tube_params": (C.c_int, [P(TubeSpec), P(Params)]),
            "orc_tube_count": (C.c_int, [P(TubeSpec), IP]),
            "orc_tube_ic": (C.c_int, [P(TubeSpec), C.c_int, P(SoaView)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _dp(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def W(r, h):
    return lib().orc_W(float(r), float(h))


def dW(r, h):
    return lib().orc_dW(float(r), float(h))


def idic_cell(vg, vd, psi, tbar, eps, tau):
    """Implicit drag solve of one averaging cell. vd: (n_dust, 3). Returns (vg_hat, vd_hat)."""
    vd = np.ascontiguousarray(vd, dtype=np.float64).reshape(-1, 3)
    nd = vd.shape[0]
    vg = np.ascontiguousarray(vg, dtype=np.float64)
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    tbar = np.ascontiguousarray(tbar, dtype=np.float64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    og = np.zeros(3)
    od = np.zeros((nd, 3))
    lib().orc_idic_cell(nd, _dp(vg), _dp(vd), _dp(psi), _dp(tbar), _dp(eps), float(tau),
                        og.ctypes.data_as(C.POINTER(C.c_double)), od.ctypes.data_as(C.POINTER(C.c_double)))
    return og, od


def tube_params(spec: TubeSpec) -> Params:
    p = Params()
    rc = lib().orc_tube_params(C.byref(spec), C.byref(p))
    if rc != OK:
        raise DshError(rc, "orc_tube_params")
    return p


def tube_count(spec: TubeSpec) -> int:
    n = (C.c_int * 3)()
    rc = lib().orc_tube_count(C.byref(spec), n)
    if rc != OK:
        raise DshError(rc, "orc_tube_count")
    return int(n[0])


def tube_ic(spec: TubeSpec, species: int) -> dict:
    arrs = empty_arrays(tube_count(spec))
    v = make_view(arrs)
    rc = lib().orc_tube_ic(C.byref(spec), species, C.byref(v))
    if rc != OK:
        raise DshError(rc, "orc_tube_ic")
    return arrs


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


UPLOAD_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "rho", "p", "e", "h", "t_stop", "ind")


class OracleSim:
    """Same operations as dustyshock3d_amd.Simulation, computed by the CPU oracle."""

    def __init__(self, params: Params):
        self._lib = lib()
        self.params = params
        self._h = C.c_void_p()
        rc = self._lib.orc_create(C.byref(params), C.byref(self._h))
        if rc != OK:
            raise DshError(rc, "orc_create")
        self.counts = {}

    @classmethod
    def from_spec(cls, spec: TubeSpec) -> "OracleSim":
        sim = cls(tube_params(spec))
        for s in (0, 1):
            sim.upload(s, tube_ic(spec, s))
        return sim

    def upload(self, species: int, arrays: dict):
        n = len(arrays["x"])
        keep = field_arrays(arrays, UPLOAD_FIELDS, n)
        v = make_view(keep)
        rc = self._lib.orc_upload(self._h, species, C.byref(v), n - 1)
        if rc != OK:
            raise DshError(rc, "orc_upload")
        self.counts[species] = n

    def step(self, nsteps: int = 1) -> float:
        t = C.c_double(0.0)
        self._lib.orc_step(self._h, nsteps, C.byref(t))
        return t.value

    def download(self, species: int, fields=ALL_FIELDS) -> dict:
        arrs = empty_arrays(self.counts[species], fields)
        v = make_view(arrs)
        rc = self._lib.orc_download(self._h, species, C.byref(v))
        if rc != OK:
            raise DshError(rc, "orc_download")
        return arrs

    def conservation(self) -> np.ndarray:
        out = np.zeros(4)
        self._lib.orc_conservation(self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def status(self) -> dict:
        t, dt = C.c_double(), C.c_double()
        steps, done = C.c_int(), C.c_int()
        self._lib.orc_status(self._h, C.byref(t), C.byref(dt), C.byref(steps), C.byref(done))
        return {"t": t.value, "dt": dt.value, "steps": steps.value, "done": bool(done.value)}

    def neighbour_stats(self, species: int):
        n = self.counts[species]
        cnt = np.zeros(n, np.int32)
        h = np.zeros(n, np.uint64)
        self._lib.orc_neighbour_stats(self._h, species, cnt.ctypes.data_as(C.POINTER(C.c_int)),
                                      h.ctypes.data_as(C.POINTER(C.c_uint64)))
        return cnt, h

    def sorted_ids(self, species: int) -> np.ndarray:
        out = np.zeros(self.counts[species], np.int32)
        rc = self._lib.orc_sorted_ids(self._h, species, out.ctypes.data_as(C.POINTER(C.c_int)))
        if rc != OK:
            raise DshError(rc, "orc_sorted_ids")
        return out

    def grid(self):
        d = np.zeros(9)
        i = np.zeros(4, np.int32)
        self._lib.orc_grid(self._h, d.ctypes.data_as(C.POINTER(C.c_double)), i.ctypes.data_as(C.POINTER(C.c_int)))
        return {"origin": d[0:3], "cell": d[3:6], "length": d[6:9], "n": i[0:3], "f": int(i[3])}

    def avg_cells(self):
        A = self._lib.orc_num_avg_cells(self._h)
        nd = self.params.n_dust
        out = {"ng": np.zeros(A, np.int32), "nd": np.zeros(A * nd, np.int32),
               "vg_hat": np.zeros(3 * A), "vd_hat": np.zeros(3 * A * nd), "psi": np.zeros(3 * A),
               "tbar": np.zeros(A * nd), "eps": np.zeros(A * nd)}
        ptr = {k: (v.ctypes.data_as(C.POINTER(C.c_int)) if v.dtype == np.int32 else
                   v.ctypes.data_as(C.POINTER(C.c_double))) for k, v in out.items()}
        rc = self._lib.orc_avg_cells(self._h, ptr["ng"], ptr["nd"], ptr["vg_hat"], ptr["vd_hat"], ptr["psi"],
                                     ptr["tbar"], ptr["eps"])
        if rc != OK:
            raise DshError(rc, "orc_avg_cells")
        return out

    def close(self):
        if self._h:
            self._lib.orc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
