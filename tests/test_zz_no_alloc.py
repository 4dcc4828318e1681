"""dsh_step_noalloc (no device allocation: its scratch is a static device array) against
dsh_step (create, upload, step, download) through the C ABI.  Two calls in a row on one GPU,
the second with a smaller tube and the other precision, so the second runs on the scratch the
first left behind: every output field, the time and the conservation sums are bit-identical
to the existing entry point's.  A tube too large for the scratch is rejected before it runs.
The test's own device arrays come from the HIP runtime libdsh.so links (hipMalloc through
ctypes), not from torch, whose wheel carries a second HIP runtime."""
import ctypes as C

import numpy as np
import pytest

from dustyshock3d_amd import Simulation, abi, configs
from dustyshock3d_amd.sim import tube_count, tube_ic, tube_params

pytestmark = pytest.mark.gpu

IN_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "rho", "p", "e", "h", "t_stop", "ind")
OUT_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "rho", "p", "e", "h", "t_stop", "edot",
              "ind", "cell")


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libdsh.so is linked against."""

    def __init__(self):
        abi.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        self.live.append(p)
        return p

    def get(self, p: C.c_void_p, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def _typed(p, field):
    return C.cast(p, C.POINTER(C.c_int if field in abi.INT_FIELDS else C.c_double))


def _noalloc(hip, params, state, counts, nsteps):
    nsp = params.n_dust + 1
    ins, outs = (abi.SoaView * nsp)(), (abi.SoaView * nsp)()
    res = {}
    for s in range(nsp):
        v = abi.SoaView()
        for f in IN_FIELDS:
            setattr(v, f, _typed(hip.put(state[s][f]), f))
        ins[s] = v
        o = abi.SoaView()
        res[s] = {}
        for f in OUT_FIELDS:
            z = np.zeros(counts[s], dtype=np.int32 if f in abi.INT_FIELDS else np.float64)
            p = hip.put(z)
            res[s][f] = (p, z)
            setattr(o, f, _typed(p, f))
        outs[s] = o
    t_out = C.c_double(-1.0)
    cons = (C.c_double * 4)()
    rc = abi.lib().dsh_step_noalloc(C.byref(params), ins, (C.c_int * nsp)(*counts), nsteps, outs, C.byref(t_out),
                                    cons, None)
    assert rc == abi.OK, rc
    got = {s: {f: hip.get(p, z) for f, (p, z) in res[s].items()} for s in range(nsp)}
    return got, t_out.value, np.array(cons[:])


def _check_against_step(spec, nsteps):
    p = tube_params(spec)
    state = {s: tube_ic(spec, s) for s in (0, 1)}
    n = tube_count(spec)
    print(f"\nno-alloc step: {n} particles per species, precision {spec.precision}, {nsteps} steps", flush=True)
    hip = _Hip()
    try:
        got, t, cons = _noalloc(hip, p, state, [n, n], nsteps)
    finally:
        hip.free()
    ref = Simulation(p)
    for s in (0, 1):
        ref.upload(s, state[s])
    ref.step(nsteps)
    assert t == ref.status()["t"]
    assert np.array_equal(cons, ref.conservation())
    for s in (0, 1):
        r = ref.download(s, OUT_FIELDS)
        for f in OUT_FIELDS:
            assert np.array_equal(got[s][f], r[f]), (s, f)
    ref.close()


def test_noalloc_matches_dsh_step_twice():
    _check_against_step(configs.cfg0(precision=abi.MIXED, max_steps=6), 6)
    _check_against_step(configs.tube(0.03, 0.01, 5.0, 0.3, max_steps=4, precision=abi.FP64, tube_index=3), 4)


def test_noalloc_rejects_a_tube_beyond_the_scratch():
    p = tube_params(configs.cfg2())
    hip = _Hip()
    try:
        v = abi.SoaView()
        for f in IN_FIELDS:
            setattr(v, f, _typed(hip.put(np.zeros(8, np.int32 if f in abi.INT_FIELDS else np.float64)), f))
        ins = (abi.SoaView * 2)(v, v)
        outs = (abi.SoaView * 2)()
        counts = (C.c_int * 2)(2812500, 2812500)   # checked before any input is read
        print("\nno-alloc step: 2 x 2812500 particles, expecting a rejection", flush=True)
        assert abi.lib().dsh_step_noalloc(C.byref(p), ins, counts, 1, outs, None, None, None) == abi.ERR_BUDGET
    finally:
        hip.free()
