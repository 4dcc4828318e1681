"""Scheduling logic of the continuous-batching server on the host-only fake device:
arrival-order admission under the HBM budget, rejection, cancellation, recovery from a
failed rebuild, finished tubes that wait for the next admission instead of forcing a
rebuild, and the time accounting the p50/p99 report uses (injected clock); the estimated
work that orders a sweep's batches lightest first."""
import math

import pytest

from dustyshock3d_amd import abi, configs
from dustyshock3d_amd.abi import DshError
from dustyshock3d_amd.batch import Server, footprint, run_sweep, sweep_groups, work
from dustyshock3d_amd.sim import tube_count


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1.0   # every clock read advances one tick
        return self.t


def small(i, steps=30):
    return configs.tube(0.03, 0.01, 1.0, 1.0, max_steps=steps, tube_index=i)


def run_until_idle(sv, limit=1000):
    for _ in range(limit):
        r, q = sv.pump()
        if r == 0 and q == 0:
            return
    raise AssertionError("server did not drain")


def test_admission_respects_budget_and_order():
    fp = footprint(small(0))
    sv = Server(hbm_budget=2.5 * fp, hbm_capacity=5 * fp, steps_per_epoch=10, backend=abi.BACKEND_FAKE,
                clock=Clock())
    ids = [sv.submit(small(i), arrival=0.0) for i in range(5)]
    r, q = sv.pump()
    assert (r, q) == (2, 3)                      # two fit the budget
    assert sv.stats()["bytes_in_use"] <= 2.5 * fp
    run_until_idle(sv)
    polls = [sv.poll(i) for i in ids]
    assert all(p["state"] == abi.REQ_DONE for p in polls)
    adm = [p["admitted"] for p in polls]
    assert adm == sorted(adm)                     # admitted in arrival order
    assert all(p["finished"] > p["admitted"] >= p["arrival"] for p in polls)
    assert all(p["particle_steps"] > 0 for p in polls)


def test_oversized_request_is_rejected_not_blocking():
    fp = footprint(small(0))
    big = configs.cfg1()
    sv = Server(hbm_budget=3 * fp, steps_per_epoch=10, backend=abi.BACKEND_FAKE, clock=Clock())
    a = sv.submit(big, 0.0)
    b = sv.submit(small(1), 0.0)
    sv.pump()
    assert sv.poll(a)["state"] == abi.REQ_REJECTED
    assert sv.poll(b)["state"] in (abi.REQ_RUNNING, abi.REQ_DONE)
    run_until_idle(sv)
    assert sv.poll(b)["state"] == abi.REQ_DONE


def test_cancel_queued_and_running():
    fp = footprint(small(0))
    sv = Server(hbm_budget=1.5 * fp, hbm_capacity=4 * fp, steps_per_epoch=5, backend=abi.BACKEND_FAKE,
                clock=Clock())
    a = sv.submit(small(0, 100), 0.0)
    b = sv.submit(small(1, 100), 0.0)
    sv.pump()
    assert sv.poll(a)["state"] == abi.REQ_RUNNING and sv.poll(b)["state"] == abi.REQ_QUEUED
    sv.cancel(b)
    sv.cancel(a)
    run_until_idle(sv)
    assert sv.poll(a)["state"] == abi.REQ_CANCELLED
    assert sv.poll(b)["state"] == abi.REQ_CANCELLED
    assert sv.stats()["bytes_in_use"] == 0


def test_failed_rebuild_keeps_running_batch_and_retries():
    fp = footprint(small(0))
    sv = Server(hbm_budget=4 * fp, hbm_capacity=8 * fp, steps_per_epoch=10, backend=abi.BACKEND_FAKE,
                clock=Clock(), fake_fail_rebuild=2)
    a = sv.submit(small(0, 50), 0.0)
    sv.pump()                                  # rebuild 1: a admitted
    b = sv.submit(small(1, 20), 5.0)
    sv.pump()                                  # rebuild 2 fails: a keeps running, b stays queued
    assert sv.poll(a)["state"] == abi.REQ_RUNNING
    assert sv.poll(b)["state"] == abi.REQ_QUEUED
    run_until_idle(sv)                         # rebuild 3 admits b
    assert sv.poll(a)["state"] == abi.REQ_DONE and sv.poll(b)["state"] == abi.REQ_DONE


def test_new_arrivals_join_a_running_batch():
    """Continuous batching: a request arriving while another runs starts before it ends."""
    fp = footprint(small(0))
    sv = Server(hbm_budget=4 * fp, hbm_capacity=8 * fp, steps_per_epoch=10, backend=abi.BACKEND_FAKE,
                clock=Clock())
    a = sv.submit(small(0, 100), 0.0)
    sv.pump()
    sv.pump()
    b = sv.submit(small(1, 20), 10.0)
    sv.pump()
    assert sv.poll(b)["state"] in (abi.REQ_RUNNING, abi.REQ_DONE)
    assert sv.poll(a)["state"] == abi.REQ_RUNNING
    run_until_idle(sv)
    pa, pb = sv.poll(a), sv.poll(b)
    assert pb["finished"] < pa["finished"]
    assert pytest.approx(pb["particle_steps"]) == 20 * 2 * 900


def test_finished_tubes_cost_no_rebuild_of_their_own():
    """A tube that finishes stays in the batch (skipped by the kernels) until requests are
    admitted; the batch is released once nothing runs."""
    fp = footprint(small(0))
    sv = Server(hbm_budget=8 * fp, hbm_capacity=16 * fp, steps_per_epoch=10, backend=abi.BACKEND_FAKE,
                clock=Clock())
    ids = [sv.submit(small(i, steps), 0.0) for i, steps in enumerate((20, 50, 90))]
    run_until_idle(sv)
    assert [sv.poll(i)["state"] for i in ids] == [abi.REQ_DONE] * 3
    assert sv.stats()["rebuilds"] == 2            # admission + release
    assert sv.stats()["bytes_in_use"] == 0
    fin = [sv.poll(i)["finished"] for i in ids]
    assert fin == sorted(fin) and len(set(fin)) == 3
    # an admission drops the finished tubes and keeps the running one
    a = sv.submit(small(3, 100), 0.0)
    sv.pump()
    b = sv.submit(small(4, 10), 0.0)
    sv.pump()
    sv.pump()
    assert sv.poll(b)["state"] == abi.REQ_DONE and sv.poll(a)["state"] == abi.REQ_RUNNING
    c = sv.submit(small(5, 10), 0.0)
    sv.pump()                                      # rebuild: a kept, b dropped, c added
    assert sv.stats()["bytes_in_use"] == pytest.approx(2 * fp)
    run_until_idle(sv)
    assert sv.poll(a)["state"] == abi.REQ_DONE and sv.poll(c)["state"] == abi.REQ_DONE


def test_heavy_requests_run_one_at_a_time_in_arrival_order():
    """With heavy_work set, requests above it (particles x estimated steps) start one at a time
    in arrival order; light requests behind a waiting heavy one wait too (FIFO), those ahead of
    it batch with the running heavy one."""
    heavy = [small(i, steps=300) for i in range(2)]
    light = [small(10 + i, steps=20) for i in range(3)]
    n = 2 * tube_count(heavy[0])
    work = n * 100          # between the two classes: n x 300 and n x 20 particle-steps
    sv = Server(hbm_budget=1e12, steps_per_epoch=10, backend=abi.BACKEND_FAKE, clock=Clock(), heavy_work=work)
    h0 = sv.submit(heavy[0], 0.0)
    l0 = sv.submit(light[0], 0.0)
    h1 = sv.submit(heavy[1], 0.0)
    l1 = sv.submit(light[1], 0.0)
    r, q = sv.pump()
    assert (r, q) == (2, 2)                      # h0 and l0 run; h1 waits, and l1 behind it
    assert sv.poll(h1)["state"] == abi.REQ_QUEUED and sv.poll(l1)["state"] == abi.REQ_QUEUED
    while sv.poll(h0)["state"] == abi.REQ_RUNNING:
        assert sv.poll(h1)["state"] == abi.REQ_QUEUED
        sv.pump()
    run_until_idle(sv)
    p = {k: sv.poll(k) for k in (h0, l0, h1, l1)}
    assert all(v["state"] == abi.REQ_DONE for v in p.values())
    assert p[h1]["admitted"] >= p[h0]["finished"]
    # without the distinction both heavy requests start at once
    sv2 = Server(hbm_budget=1e12, steps_per_epoch=10, backend=abi.BACKEND_FAKE, clock=Clock())
    a, b = sv2.submit(heavy[0], 0.0), sv2.submit(heavy[1], 0.0)
    assert sv2.pump() == (2, 0)


def test_sweep_groups_run_lightest_half_first():
    # uneven tubes: work = particles x steps, the steps fixed by max_steps
    specs = [configs.tube(w, dx, 1.0, 1.0, max_steps=st, tube_index=i)
             for i, (w, dx, st) in enumerate([(0.05, 0.005, 40), (0.03, 0.01, 10), (0.04, 0.006, 25),
                                              (0.03, 0.01, 30), (0.05, 0.008, 5), (0.02, 0.01, 20),
                                              (0.04, 0.005, 10)])]
    w = [work(sp) for sp in specs]
    for sp, wk in zip(specs, w):
        assert wk == 2 * tube_count(sp) * sp.max_steps
    assert sweep_groups(specs, 1) == [list(range(len(specs)))]
    g2 = sweep_groups(specs, 2)
    assert [len(g) for g in g2] == [3, 4]
    g3 = sweep_groups(specs, 3)
    assert [len(g) for g in g3] == [3, 2, 2]
    for groups in (g2, g3):
        flat = [i for g in groups for i in g]
        assert sorted(flat) == list(range(len(specs)))
        # ascending work along the batches, ties in index order
        assert [w[i] for i in flat] == sorted(w)
        assert all(w[a] < w[b] or (w[a] == w[b] and a < b) for a, b in zip(flat, flat[1:]))


def test_work_of_a_timed_tube_counts_cfl_steps_at_rest():
    # t_end over the CFL step of the set-up at rest, cfl x 1.2 dx (unit sound speed)
    sp = configs.tube(0.03, 0.01, 1.0, 1.0, t_end=0.2, dt_mode=abi.DT_CFL, cfl=0.3)
    assert work(sp) == 2 * tube_count(sp) * math.ceil(0.2 / (0.3 * 1.2 * 0.01))
    with pytest.raises(DshError):
        work(configs.tube(0.03, -0.01, 1.0, 1.0, max_steps=5))


def test_sweep_rejects_a_tube_without_an_end():
    # checked before any batch is built, so no device is needed
    with pytest.raises(ValueError):
        run_sweep([configs.tube(0.03, 0.01, 1.0, 1.0, max_steps=5), configs.tube(0.03, 0.01, 1.0, 1.0)])
