"""Batching and serving invariance on the GPU: a tube's result does not depend on the
other tubes in its batch, nor on being moved between batches by the server.  Every
per-tube reduction runs in a fixed order inside the tube, so the results are bit-identical."""
import numpy as np
import pytest

from dustyshock3d_amd import Simulation, abi, configs
from dustyshock3d_amd.batch import Batch, Server

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "rho", "h", "e", "t_stop", "ax", "ay", "az")


def small_specs(precision):
    # uneven sizes, drag and dust loading, from the sweep distribution but short tubes
    out = []
    for i, (w, dx, K, d2g) in enumerate([(0.03, 0.01, 0.05, 0.3), (0.05, 0.006, 200.0, 2.0),
                                         (0.02, 0.008, 3.0, 8.0), (0.04, 0.005, 1000.0, 0.05)]):
        out.append(configs.tube(w, dx, K, d2g, t_end=0.02, dt_mode=abi.DT_CFL, precision=precision, tube_index=i))
    return out


def alone(spec, steps):
    s = Simulation.from_spec(spec)
    s.step(steps)
    return s


@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_batch_matches_alone_bit_exact(precision):
    specs = small_specs(precision)
    print(f"\nbatch precision={precision} tubes={len(specs)}", flush=True)
    b = Batch(specs)
    b.run(400)
    for i, spec in enumerate(specs):
        st = b.status(i)
        assert st["done"] and abs(st["t"] - 0.02) < 1e-12
        s = alone(spec, 400)
        assert s.status()["steps"] == st["steps"]
        for sp in (0, 1):
            rb, rs = b.download(i, sp), s.download(sp)
            for k in FIELDS:
                assert np.array_equal(rb[k], rs[k]), (i, sp, k)
        assert np.array_equal(b.conservation(i), s.conservation())


def test_server_results_match_alone():
    specs = small_specs(abi.MIXED)
    sv = Server(hbm_budget=64e9, steps_per_epoch=7)
    ids = [sv.submit(specs[0], 0.0), sv.submit(specs[1], 0.0)]
    sv.pump()
    sv.pump()
    ids += [sv.submit(specs[2], 1.0), sv.submit(specs[3], 1.0)]   # join a running batch
    for _ in range(200):
        r, q = sv.pump()
        if r == 0 and q == 0:
            break
    st = sv.stats()
    print(f"\nserver rebuilds={st['rebuilds']} epochs={st['epochs']}", flush=True)
    assert st["rebuilds"] >= 2
    for rid, spec in zip(ids, specs):
        p = sv.poll(rid)
        assert p["state"] == abi.REQ_DONE
        s = alone(spec, 400)
        for sp in (0, 1):
            rv, rs = sv.result(rid, sp), s.download(sp)
            for k in FIELDS:
                assert np.array_equal(rv[k], rs[k]), (rid, sp, k)


def _raw_batch(arrays_per_tube, params):
    import ctypes as C
    from dustyshock3d_amd.abi import Params, make_view
    lib = abi.lib()
    h = C.c_void_p()
    assert lib.dsh_batch_create((Params * len(params))(*params), len(params), C.byref(h)) == abi.OK
    for t, per_species in enumerate(arrays_per_tube):
        for s, a in enumerate(per_species):
            v = make_view({k: a[k] for k in ("x", "y", "z", "vx", "vy", "vz", "m", "rho", "p", "e", "h",
                                             "t_stop", "ind")})
            assert lib.dsh_batch_upload(h, t, s, C.byref(v), len(a["x"]) - 1) == abi.OK
    return lib, h


def test_failed_tube_fails_alone():
    """A tube whose smoothing length outgrows its grid's search window stops with an error
    before any pair kernel runs for it; the other tube of the batch completes."""
    import ctypes as C
    from dustyshock3d_amd.sim import tube_ic, tube_params
    good = configs.tube(0.03, 0.01, 1.0, 1.0, max_steps=5, tube_index=1)
    bad = configs.tube(0.03, 0.01, 1.0, 1.0, max_steps=5, tube_index=2)
    arrays = [[tube_ic(good, 0), tube_ic(good, 1)], [tube_ic(bad, 0), tube_ic(bad, 1)]]
    # the grid comes from the smallest initial h; one particle with 200x that h needs a
    # search window of ~190 cells, above the 64-cell limit
    i = len(arrays[1][0]["h"]) // 2
    arrays[1][0]["h"][i] = 200.0 * arrays[1][0]["h"].min()
    lib, h = _raw_batch(arrays, [tube_params(good), tube_params(bad)])
    nd = C.c_int()
    assert lib.dsh_batch_run(h, 5, C.byref(nd)) == abi.ERR_GRID
    assert b"tube 1" in lib.dsh_batch_last_error(h)
    steps, done = C.c_int(), C.c_int()
    lib.dsh_batch_status(h, 0, None, None, C.byref(steps), C.byref(done))
    assert (steps.value, done.value) == (5, 1)
    lib.dsh_batch_status(h, 1, None, None, C.byref(steps), C.byref(done))
    assert (steps.value, done.value) == (0, 1)
    lib.dsh_batch_destroy(h)


def test_device_setup_matches_host_setup():
    """Batches generate their tubes on the device (k_tube_ic); before any step the state
    equals the host set-up dsh_tube_ic (itself bit-exact against the oracle) bit for bit."""
    from dustyshock3d_amd import sim
    specs = [configs.cfg0(precision=abi.MIXED), configs.sweep_tube(3), configs.sweep_tube(11), configs.cfg1()]
    b = Batch(specs)
    for i, spec in enumerate(specs):
        for sp in (0, 1):
            host = sim.tube_ic(spec, sp)
            dev = b.download(i, sp)
            assert len(dev["x"]) == len(host["x"])
            for k in ("x", "y", "z", "vx", "vy", "vz", "m", "rho", "p", "e", "h", "t_stop", "ind"):
                assert np.array_equal(dev[k], host[k]), (i, sp, k)


def test_mixed_physics_batch_matches_alone():
    """One batch holding a gas-only tube (plain scheme), a K-law dusty tube and a tube with the
    reference's own stopping-time law and two dust species: each equals its solo run bit for bit."""
    import ctypes as C
    import os
    from dustyshock3d_amd import sim
    from tests.parity_util import inject
    gas_spec = configs.tube(0.03, 0.01, 1.0, 1.0, max_steps=6, tube_index=3)
    p_gas = sim.tube_params(gas_spec)
    p_gas.n_dust = 0
    st_gas = [sim.tube_ic(gas_spec, 0)]
    dusty_spec = configs.tube(0.02, 0.01, 30.0, 2.0, max_steps=6, tube_index=4)
    p_dusty = sim.tube_params(dusty_spec)
    st_dusty = [sim.tube_ic(dusty_spec, 0), sim.tube_ic(dusty_spec, 1)]
    p_ref = sim.read_init(os.path.join(os.path.dirname(__file__), "golden", "Init.txt"))
    p_ref.particles_per_length, p_ref.h, p_ref.max_steps = 20, 0.2, 6
    sim.reference_setup(p_ref)
    st_ref = [sim.reference_ic(p_ref, s) for s in range(3)]
    params = [p_gas, p_dusty, p_ref]
    states = [st_gas, st_dusty, st_ref]
    lib, h = _raw_batch(states, params)
    nd = C.c_int()
    assert lib.dsh_batch_run(h, 50, C.byref(nd)) == abi.OK
    assert nd.value == 3
    for t, (p, st) in enumerate(zip(params, states)):
        solo = inject(Simulation, p, {s: a for s, a in enumerate(st)})
        solo.step(50)
        for s in range(len(st)):
            arrs = abi.empty_arrays(len(st[s]["x"]))
            v = abi.make_view(arrs)
            assert lib.dsh_batch_download(h, t, s, C.byref(v)) == abi.OK
            ref = solo.download(s)
            for k in FIELDS:
                assert np.array_equal(arrs[k], ref[k]), (t, s, k)
    lib.dsh_batch_destroy(h)


@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_graph_replays_match_eager_launches(precision):
    """Steps replayed from the captured graphs (eight steps per launch while eight are left,
    then single steps) equal steps launched one kernel at a time, bit for bit, with eager
    work and synchronisations between the replays."""
    spec = configs.cfg0(precision=precision, max_steps=40)
    g = Simulation.from_spec(spec)
    e = Simulation.from_spec(spec)
    e.set_timing(True)          # stage events: every kernel launched eagerly
    for n in (19, 1, 8, 5):     # 2 x 8 + 3, 1, 8, 5: both graphs, several replays each
        print(f"\ngraph vs eager precision={precision} steps={n}", flush=True)
        g.step(n)
        e.step(n)
        assert g.status() == e.status()
        for s in (0, 1):
            dg, de = g.download(s), e.download(s)
            for k in FIELDS:
                assert np.array_equal(dg[k], de[k]), (n, s, k)
        assert np.array_equal(g.conservation(), e.conservation())


@pytest.mark.parametrize("concurrent", [False, True])
def test_sweep_lightest_half_first_matches_one_batch(concurrent):
    """run_sweep in two batches (the lighter half by estimated work, and the rest), one after
    the other or side by side on their own streams and threads, gives every tube the result
    of one batch of all tubes, bit for bit."""
    from dustyshock3d_amd.batch import run_sweep, sweep_groups

    specs = small_specs(abi.MIXED)
    print(f"\nsweep of {len(specs)} tubes in two batches, concurrent={concurrent}", flush=True)
    got = {}

    def keep(i, b, j):
        got[i] = ([b.download(j, sp) for sp in (0, 1)], b.conservation(j))

    r = run_sweep(specs, batches=2, chunk=5, on_done=keep, concurrent=concurrent)
    assert sorted(got) == list(range(len(specs)))
    light, heavy = sweep_groups(specs, 2)
    if not concurrent:
        assert max(r["finish_s"][i] for i in light) < min(r["finish_s"][i] for i in heavy)
    ref = Batch(specs)
    ref.run(400)
    for i in range(len(specs)):
        assert ref.status(i)["done"] and ref.status(i)["steps"] == r["steps"][i]
        for sp in (0, 1):
            rr = ref.download(i, sp)
            for k in FIELDS:
                assert np.array_equal(got[i][0][sp][k], rr[k]), (i, sp, k)
        assert np.array_equal(got[i][1], ref.conservation(i))
