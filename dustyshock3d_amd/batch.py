"""Batches of independent tubes and the continuous-batching server, over the C ABI."""
from __future__ import annotations

import ctypes as C
import threading
import time
from typing import Callable, Optional, Sequence

import numpy as np

from . import abi
from .abi import DshError, ServerOpts, TubeSpec, empty_arrays, make_view
from .sim import tube_count


def _times(fn, handle) -> dict:
    nmax, ln = 32, 32
    ms = np.zeros(nmax)
    names = C.create_string_buffer(nmax * ln)
    n = fn(handle, ms.ctypes.data_as(C.POINTER(C.c_double)), names, nmax, ln)
    raw = names.raw
    return {raw[i * ln:(i + 1) * ln].split(b"\0")[0].decode(): float(ms[i]) for i in range(min(max(n, 0), nmax))}


class Batch:
    """Many tubes stepped together: one launch sequence per step for the whole batch."""

    def __init__(self, specs: Sequence[TubeSpec]):
        self._lib = abi.lib()
        self.specs = list(specs)
        arr = (TubeSpec * len(self.specs))(*self.specs)
        self._h = C.c_void_p()
        rc = self._lib.dsh_batch_from_specs(arr, len(self.specs), C.byref(self._h))
        if rc != abi.OK:
            raise DshError(rc, "dsh_batch_from_specs")
        self.counts = [tube_count(s) for s in self.specs]

    def _call(self, rc):
        if rc != abi.OK:
            m = self._lib.dsh_batch_last_error(self._h)
            raise DshError(rc, m.decode() if m else "")

    def run(self, nsteps: int) -> int:
        nd = C.c_int(0)
        self._call(self._lib.dsh_batch_run(self._h, nsteps, C.byref(nd)))
        return nd.value

    def status(self, tube: int) -> dict:
        t, dt = C.c_double(), C.c_double()
        steps, done = C.c_int(), C.c_int()
        self._call(self._lib.dsh_batch_status(self._h, tube, C.byref(t), C.byref(dt), C.byref(steps), C.byref(done)))
        return {"t": t.value, "dt": dt.value, "steps": steps.value, "done": bool(done.value)}

    def download(self, tube: int, species: int, fields=abi.ALL_FIELDS) -> dict:
        arrs = empty_arrays(self.counts[tube], fields)
        v = make_view(arrs)
        self._call(self._lib.dsh_batch_download(self._h, tube, species, C.byref(v)))
        return arrs

    def conservation(self, tube: int) -> np.ndarray:
        out = np.zeros(4)
        self._call(self._lib.dsh_batch_conservation(self._h, tube, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set_tile(self, tile: int):
        """Candidates staged per LDS piece by the tiled pair kernels (128, 160, 192, 224, 256)."""
        self._call(self._lib.dsh_batch_set_tile(self._h, int(tile)))

    def set_block_waves(self, waves: int):
        """Waves (64-target chunks) per block of the tiled pair kernels (1, 2, 4)."""
        self._call(self._lib.dsh_batch_set_block_waves(self._h, int(waves)))

    def block_shape(self) -> dict:
        shape = (C.c_int * 2)()
        self._call(self._lib.dsh_batch_block_shape(self._h, shape))
        return {"tile": int(shape[0]), "waves": int(shape[1])}

    def autotune(self, steps_per_trial: int = 2, with_variants: bool = False) -> dict:
        """Time each pair-kernel configuration (piece size x waves per block) on the live
        state and keep the fastest."""
        chosen = (C.c_int * 2)()
        ms = C.c_double(0.0)
        self._call(self._lib.dsh_batch_autotune(self._h, int(steps_per_trial), int(with_variants), chosen, C.byref(ms)))
        return {"tile": int(chosen[0]), "waves": self.block_shape()["waves"], "variant": int(chosen[1]),
                "ms_per_step": ms.value}

    def set_pair_variant(self, v: int):
        self._call(self._lib.dsh_batch_set_pair_variant(self._h, int(v)))

    def set_timing(self, on: bool = True):
        self._call(self._lib.dsh_batch_set_timing(self._h, int(on)))

    def kernel_times(self) -> dict:
        return _times(self._lib.dsh_batch_kernel_times, self._h)

    def device_bytes(self) -> float:
        b = C.c_double()
        self._call(self._lib.dsh_batch_device_bytes(self._h, C.byref(b)))
        return b.value

    def close(self):
        if self._h:
            self._lib.dsh_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def footprint(spec: TubeSpec) -> float:
    b = C.c_double()
    rc = abi.lib().dsh_server_footprint(C.byref(spec), C.byref(b))
    if rc != abi.OK:
        raise DshError(rc, "dsh_server_footprint")
    return b.value


def work(spec: TubeSpec) -> float:
    """Estimated work of a tube: particles x estimated steps (the server's cost estimate)."""
    w = C.c_double()
    rc = abi.lib().dsh_server_work(C.byref(spec), C.byref(w))
    if rc != abi.OK:
        raise DshError(rc, "dsh_server_work")
    return w.value


def sweep_groups(specs: Sequence[TubeSpec], batches: int = 2) -> list:
    """Tube indices of a sweep's batches, lightest first.

    The tubes are ordered by estimated work (ties by index); each batch but the last takes the
    lighter half of the tubes not yet placed, the last takes the rest.  batches = 1: one batch
    of all tubes in index order."""
    n = len(specs)
    if batches <= 1 or n < 2:
        return [list(range(n))]
    w = np.array([work(sp) for sp in specs])
    order = [int(i) for i in np.argsort(w, kind="stable")]
    groups = []
    while len(groups) < batches - 1 and len(order) >= 2:
        k = len(order) // 2
        groups.append(order[:k])
        order = order[k:]
    groups.append(order)
    return groups


def run_sweep(specs: Sequence[TubeSpec], batches: int = 1, chunk: int = 16,
              pair_variant: int = abi.PAIR_TILED, on_done: Optional[Callable] = None,
              clock: Callable[[], float] = time.perf_counter, concurrent: bool = True) -> dict:
    """Run every tube of a sweep to its end in the batches of sweep_groups.

    concurrent: every batch has its own stream and host thread, so the batches share the GPU.
    In one batch every tube advances one step per batch step, so a light tube's time to
    solution is its steps times the step of the whole batch; in a batch of its own, beside the
    heavy one, it runs at its own batch's pace.  Otherwise the batches run one after another.

    Returns per tube (index order) the time to solution in seconds from the call, batch
    set-ups included, and its steps.  on_done(i, batch, slot) is called once per tube when it
    has finished, while its batch is alive (from that batch's thread when concurrent).  A
    tube's results do not depend on the batch it runs in, nor on what runs beside it."""
    for i, sp in enumerate(specs):
        if not (sp.t_end < 1e29 or sp.max_steps > 0):
            raise ValueError(f"tube {i} has no end (t_end or max_steps)")
    t0 = clock()
    n = len(specs)
    finish, steps = [0.0] * n, [0] * n
    groups = sweep_groups(specs, batches)

    def make(g):
        b = Batch([specs[i] for i in g])
        b.set_pair_variant(pair_variant)
        b.run(0)
        return b

    def drive(g, b):
        left = list(range(len(g)))
        while left:
            b.run(chunk)
            now = clock() - t0
            still = []
            for j in left:
                st = b.status(j)
                if st["done"]:
                    finish[g[j]] = now
                    steps[g[j]] = st["steps"]
                    if on_done is not None:
                        on_done(g[j], b, j)
                else:
                    still.append(j)
            left = still

    if concurrent and len(groups) > 1:
        made = [make(g) for g in groups]
        errors = []

        def body(g, b):
            try:
                drive(g, b)
            except BaseException as e:  # re-raised in the caller's thread
                errors.append(e)

        threads = [threading.Thread(target=body, args=(g, b)) for g, b in zip(groups, made)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for b in made:
            b.close()
        if errors:
            raise errors[0]
    else:
        for g in groups:
            b = make(g)
            drive(g, b)
            b.close()
    return {"finish_s": finish, "steps": steps}


class Server:
    """Continuous batching of tube requests (GPU engine or the host-only fake device)."""

    def __init__(self, hbm_budget: float, steps_per_epoch: int = 10, backend: int = abi.BACKEND_GPU,
                 clock: Optional[Callable[[], float]] = None, hbm_capacity: float = 0.0,
                 pair_variant: int = abi.PAIR_TILED, fake_fail_rebuild: int = 0, heavy_work: float = 0.0,
                 max_heavy: int = 1):
        self._lib = abi.lib()
        self._clock_cb = abi.CLOCK_FN(lambda _u: float(clock())) if clock else abi.CLOCK_FN()
        o = ServerOpts(hbm_budget=hbm_budget, hbm_capacity=hbm_capacity, steps_per_epoch=steps_per_epoch,
                       backend=backend, pair_variant=pair_variant, clock=self._clock_cb, clock_user=None,
                       fake_fail_rebuild=fake_fail_rebuild, heavy_work=heavy_work, max_heavy=max_heavy)
        self._h = C.c_void_p()
        rc = self._lib.dsh_server_create(C.byref(o), C.byref(self._h))
        if rc != abi.OK:
            raise DshError(rc, "dsh_server_create")
        self.specs = {}

    def _call(self, rc):
        if rc != abi.OK:
            m = self._lib.dsh_server_last_error(self._h)
            raise DshError(rc, m.decode() if m else "")

    def submit(self, spec: TubeSpec, arrival: float) -> int:
        i = C.c_uint64()
        self._call(self._lib.dsh_server_submit(self._h, C.byref(spec), float(arrival), C.byref(i)))
        self.specs[i.value] = spec
        return i.value

    def pump(self):
        r, q = C.c_int(), C.c_int()
        self._call(self._lib.dsh_server_pump(self._h, C.byref(r), C.byref(q)))
        return r.value, q.value

    def poll(self, rid: int) -> dict:
        st = C.c_int()
        a, ad, f, ps = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self._call(self._lib.dsh_server_poll(self._h, rid, C.byref(st), C.byref(a), C.byref(ad), C.byref(f),
                                             C.byref(ps)))
        return {"state": st.value, "arrival": a.value, "admitted": ad.value, "finished": f.value,
                "particle_steps": ps.value}

    def result(self, rid: int, species: int, fields=abi.DOUBLE_FIELDS + ("ind",)) -> dict:
        arrs = empty_arrays(tube_count(self.specs[rid]), fields)
        v = make_view(arrs)
        self._call(self._lib.dsh_server_result(self._h, rid, species, C.byref(v)))
        return arrs

    def cancel(self, rid: int):
        self._call(self._lib.dsh_server_cancel(self._h, rid))

    def stats(self) -> dict:
        b = C.c_double()
        r, e = C.c_int(), C.c_int()
        self._call(self._lib.dsh_server_stats(self._h, C.byref(b), C.byref(r), C.byref(e)))
        rs, adv = C.c_double(), C.c_double()
        self._call(self._lib.dsh_server_times(self._h, C.byref(rs), C.byref(adv)))
        return {"bytes_in_use": b.value, "rebuilds": r.value, "epochs": e.value,
                "rebuild_s": rs.value, "advance_s": adv.value}

    def close(self):
        if self._h:
            self._lib.dsh_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
