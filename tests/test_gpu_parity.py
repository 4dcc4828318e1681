"""GPU (HIP through the C ABI) against the FP64 oracle.

Bit-exact on a shared state: neighbour-cell keys, the sorted (cell, id) order as the
reference's Nn linked lists (kernel.cu:1937-1995) and the neighbour sets (count and an
order-free hash of (id, periodic image) per particle).  Fields within tolerance:
  FP64 mode   density, acceleration, velocities ~1e-10 (summation order only)
  mixed mode  density <= 1e-5 pointwise, velocities <= 1e-5 normwise (FP32 pair terms)
  momentum    |dP| / max(|P|, sum m|v|) <= 1e-9 in both modes, also after 50 mixed steps.
              A tube's momentum changes only through its frozen walls: the FP32 pair terms
              between active particles are antisymmetric bit for bit and accumulate in FP64,
              so they cancel; pairs with a frozen particle, or with both particles in the FP64
              wall band, are FP64; and the band covers every particle whose FP32 rounding could
              reach a wall before the run ends (f x 0.5 cs x time left + (3 + f) cells, f the
              averaging factor), so the wall force follows the FP64 run.
  energy      |dE| / |E| <= 1e-12 (FP64), <= 1e-8 (mixed).
Over long mixed runs a particle can land on the other side of an averaging-cell face than in
the FP64 run (the cell is a truncated FP64 divide, kernel.cu:2071); its cell means, and so its
velocity, then jump.  After 50 mixed steps the velocities are therefore held at the 99.9%
quantile, and the largest difference is printed.
"""
import numpy as np
import pytest

from dustyshock3d_amd import Simulation, configs
from dustyshock3d_amd.abi import FP64, MIXED, PAIR_DIRECT, PAIR_TILED
from tests.parity_util import (energy_error, inject, momentum_error, quantile_normwise, rel_normwise, rel_pointwise,
                               take_state, vec_normwise)

pytestmark = pytest.mark.gpu


def _oracle_state(oracle_mod, spec, steps):
    o = oracle_mod.OracleSim.from_spec(spec)
    if steps:
        o.step(steps)
    return o.params, take_state(o)


def _one_step_both(oracle_mod, params, state, precision, variant=PAIR_TILED):
    params.precision = precision
    o = inject(oracle_mod.OracleSim, params, state)
    g = inject(Simulation, params, state)
    g.set_pair_variant(variant)
    g.set_stats(True)
    o.step(1)
    g.step(1)
    return o, g


@pytest.mark.parametrize("variant", [PAIR_TILED, PAIR_DIRECT])
@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("warm_steps", [0, 20])
def test_step_parity_shared_state(oracle_mod, precision, warm_steps, variant):
    spec = configs.cfg0()
    params, state = _oracle_state(oracle_mod, spec, warm_steps)
    print(f"\ncase precision={precision} warm={warm_steps} variant={variant}", flush=True)
    o, g = _one_step_both(oracle_mod, params, state, precision, variant)
    for s in (0, 1):
        ro, rg = o.download(s), g.download(s)
        # integer work: bit-exact
        assert np.array_equal(rg["cell"], ro["cell"]), f"species {s}: cell keys differ"
        assert np.array_equal(rg["nn"], ro["nn"]), f"species {s}: sorted order (Nn lists) differs"
        co, ho = o.neighbour_stats(s)
        cg, hg = g.neighbour_stats(s)
        assert np.array_equal(cg, co), f"species {s}: neighbour counts differ"
        assert np.array_equal(hg, ho), f"species {s}: neighbour sets differ"
        act = ro["ind"] == 0
        tol_rho = 1e-12 if precision == FP64 else 1e-5
        e_rho = rel_pointwise(rg["rho"][act], ro["rho"][act])
        tol_v = 1e-10 if precision == FP64 else 1e-5
        e_v = vec_normwise(rg, ro)
        e_x = max(float(np.max(np.abs(rg[k] - ro[k]))) for k in ("x", "y", "z"))
        vmax = max(float(np.max(np.abs(ro[k]))) for k in ("vx", "vy", "vz"))
        tol_x = tol_v * vmax * o.status()["dt"] + 1e-15
        print(f"species {s}: rho {e_rho:.3e} v {e_v:.3e} x {e_x:.3e} (tol {tol_x:.1e})")
        assert e_rho <= tol_rho
        assert e_v <= tol_v
        assert e_x <= tol_x
        if s == 0:
            tol_a = 1e-10 if precision == FP64 else 1e-5
            e_a = vec_normwise(rg, ro, ("ax", "ay", "az"))
            e_ed = rel_normwise(rg["edot"], ro["edot"])
            print(f"gas A {e_a:.3e} edot {e_ed:.3e}")
            assert e_a <= tol_a
            assert e_ed <= tol_a
        else:
            e_ts = rel_pointwise(rg["t_stop"][act], ro["t_stop"][act])
            assert e_ts <= tol_rho
    st_o = take_state(o)
    e_p = momentum_error(g.conservation(), o.conservation(), st_o)
    e_e = energy_error(g.conservation(), o.conservation())
    print(f"momentum {e_p:.3e} energy {e_e:.3e}")
    assert e_p <= 1e-9
    assert e_e <= (1e-12 if precision == FP64 else 1e-8)
    so, sg = o.status(), g.status()
    assert sg["steps"] == so["steps"] == 1
    assert abs(sg["t"] - so["t"]) <= 1e-15
    assert abs(sg["dt"] - so["dt"]) <= 1e-15 * so["dt"]


def test_cfg0_run_fp64(oracle_mod):
    """cfg0 end to end (50 fixed CFL steps, FP64) against the oracle."""
    spec = configs.cfg0()
    o = oracle_mod.OracleSim.from_spec(spec)
    g = Simulation.from_spec(spec)
    o.step(50)
    g.step(50)
    so, sg = o.status(), g.status()
    assert sg["steps"] == so["steps"] == 50 and sg["done"] and so["done"]
    st_o = take_state(o)
    for s in (0, 1):
        ro, rg = st_o[s], g.download(s)
        act = ro["ind"] == 0
        e_rho = rel_pointwise(rg["rho"][act], ro["rho"][act])
        e_v = vec_normwise(rg, ro)
        print(f"\ncfg0 species {s}: rho {e_rho:.3e} v {e_v:.3e}")
        assert e_rho <= 1e-5
        assert e_v <= 1e-5
    e_p = momentum_error(g.conservation(), o.conservation(), st_o)
    e_e = energy_error(g.conservation(), o.conservation())
    print(f"momentum {e_p:.3e} energy {e_e:.3e}")
    assert e_p <= 1e-9
    assert e_e <= 1e-12


def test_cfg0_mixed_run_50_steps(oracle_mod):
    """The cfg0 tube in mixed precision over its 50 steps, in lockstep with the FP64 oracle.

    Until the first particle sits in another averaging cell than in the oracle's run (a
    truncated FP64 divide, or a periodic wrap, decided differently by a rounding difference),
    every step holds density, velocity <= 1e-5 and energy <= 1e-8; that horizon must reach 10
    steps (where it ends is rounding luck; after a flip the particle's cell means, and so its
    drag heating, differ).  Momentum <= 1e-9 holds for the first 25 steps: FP32 rounding outside
    the FP64 wall band then reaches the wall layers (with the default band speed 0.5 the
    momentum difference grows to ~1e-9 by step 33 and ~2e-8 by step 50; band speed 2 keeps it at
    4e-11), so after step 25 the bound is 1e-7.  The step itself stays exact: from
    the GPU's own state at steps 25 and 50, one GPU step and one oracle step agree on sets
    (bit-exact), momentum <= 1e-9, energy <= 1e-8 and fields <= 1e-5."""
    spec = configs.cfg0(precision=MIXED)
    o = oracle_mod.OracleSim.from_spec(spec)
    g = Simulation.from_spec(spec)
    first_flip = None
    for k in range(1, 51):
        o.step(1)
        g.step(1)
        st_o = take_state(o)
        st_g = take_state(g)
        flipped = any(not np.array_equal(st_g[s]["cell"], st_o[s]["cell"]) for s in (0, 1))
        if flipped and first_flip is None:
            first_flip = k
        e_p = momentum_error(g.conservation(), o.conservation(), st_o)
        e_e = energy_error(g.conservation(), o.conservation())
        e_f = 0.0
        for s in (0, 1):
            act = st_o[s]["ind"] == 0
            e_f = max(e_f, rel_pointwise(st_g[s]["rho"][act], st_o[s]["rho"][act]), vec_normwise(st_g[s], st_o[s]))
        if first_flip is None:
            assert e_f <= 1e-5 and e_e <= 1e-8, (k, e_f, e_e)
        assert e_p <= (1e-9 if k <= 25 else 1e-7), (k, e_p)
        if k in (10, 20, 30, 40, 50) or k == first_flip:
            print(f"\nstep {k}: momentum {e_p:.3e} energy {e_e:.3e} fields {e_f:.3e}"
                  f"{' (first averaging-cell flip)' if k == first_flip else ''}", flush=True)
        if k in (25, 50):
            # one step from the GPU's own state: the arithmetic of the step itself
            params = g.params
            oo = inject(oracle_mod.OracleSim, params, st_g)
            gg = inject(Simulation, params, st_g)
            gg.set_stats(True)
            oo.step(1)
            gg.step(1)
            st_oo = take_state(oo)
            for s in (0, 1):
                ro, rg = st_oo[s], gg.download(s)
                assert np.array_equal(rg["cell"], ro["cell"]) and np.array_equal(rg["nn"], ro["nn"])
                co, ho = oo.neighbour_stats(s)
                cg, hg = gg.neighbour_stats(s)
                assert np.array_equal(cg, co) and np.array_equal(hg, ho)
                act = ro["ind"] == 0
                assert rel_pointwise(rg["rho"][act], ro["rho"][act]) <= 1e-5
                assert vec_normwise(rg, ro) <= 1e-5
            e_p1 = momentum_error(gg.conservation(), oo.conservation(), st_oo)
            e_e1 = energy_error(gg.conservation(), oo.conservation())
            print(f"one step from the GPU state at step {k}: momentum {e_p1:.3e} energy {e_e1:.3e}", flush=True)
            assert e_p1 <= 1e-9 and e_e1 <= 1e-8
    print(f"first averaging-cell flip: {first_flip}")
    assert first_flip is None or first_flip > 10


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("avg_factor", [2, 3])
def test_wide_averaging_cells_parity(oracle_mod, precision, avg_factor):
    """Averaging cells of f x f x f neighbour cells (f >= 2, one wave per averaging cell):
    10 steps of the cfg0 tube against the oracle, which restates the same grouping."""
    spec = configs.cfg0(max_steps=10)
    o0 = oracle_mod.OracleSim.from_spec(spec)
    params, state = o0.params, take_state(o0)
    params.avg_factor = avg_factor
    params.precision = precision
    print(f"\ncase avg_factor={avg_factor} precision={precision}", flush=True)
    o = inject(oracle_mod.OracleSim, params, state)
    g = inject(Simulation, params, state)
    o.step(10)
    g.step(10)
    so, sg = o.status(), g.status()
    assert sg["steps"] == so["steps"] == 10
    st_o = take_state(o)
    tol = 1e-9 if precision == FP64 else 1e-5
    for s in (0, 1):
        ro, rg = st_o[s], g.download(s)
        assert np.array_equal(rg["cell"], ro["cell"])
        act = ro["ind"] == 0
        e_rho = rel_pointwise(rg["rho"][act], ro["rho"][act])
        e_v = vec_normwise(rg, ro)
        print(f"species {s}: rho {e_rho:.3e} v {e_v:.3e}")
        assert e_rho <= tol
        assert e_v <= tol
    e_p = momentum_error(g.conservation(), o.conservation(), st_o)
    e_e = energy_error(g.conservation(), o.conservation())
    print(f"momentum {e_p:.3e} energy {e_e:.3e}")
    assert e_p <= 1e-9
    assert e_e <= (1e-12 if precision == FP64 else 1e-8)


def test_neighbours_many_periodic_images_away(oracle_mod):
    """A tube 0.001 wide with h = 0.006 in the shared state: neighbours sit up to 12 periodic
    images away in y and z, beyond the range of the FP32 force pass's fixed-point y, z
    differences (2 box lengths), so those waves take dy, dz from FP64 positions.  One step in
    mixed precision: sets bit-exact, fields within the mixed tolerances."""
    width, h = 0.001, 0.006
    spec = configs.tube(width, 0.01, 10.0, 1.0, max_steps=1, precision=MIXED)
    o0 = oracle_mod.OracleSim.from_spec(spec)
    params, state = o0.params, take_state(o0)
    for s in (0, 1):
        state[s]["h"] = np.full_like(state[s]["h"], h)
    assert 2 * h >= 4 * width
    o, g = _one_step_both(oracle_mod, params, state, MIXED)
    for s in (0, 1):
        ro, rg = o.download(s), g.download(s)
        co, ho = o.neighbour_stats(s)
        cg, hg = g.neighbour_stats(s)
        assert np.array_equal(cg, co) and np.array_equal(hg, ho)
        act = ro["ind"] == 0
        e_rho = rel_pointwise(rg["rho"][act], ro["rho"][act])
        e_v = vec_normwise(rg, ro)
        print(f"\nimages species {s}: neighbours {float(cg.mean()):.0f} per particle, rho {e_rho:.3e} v {e_v:.3e}")
        assert e_rho <= 1e-5
        assert e_v <= 1e-5
        if s == 0:
            e_a = vec_normwise(rg, ro, ("ax", "ay", "az"))
            print(f"gas A {e_a:.3e}")
            assert e_a <= 1e-5


def test_dense_cells_parity(oracle_mod):
    """The cfg0 tube with h = 0.025 in the shared state: cells of ~0.05 hold up to ~1060
    members (the rank pass ranks from memory instead of LDS for runs over 1024) and each
    particle has ~3600 neighbours (the force pass runs out of stored hit masks and falls back
    to its own predicate).  One step in mixed precision: keys, order and sets bit-exact,
    fields within the mixed tolerances."""
    spec = configs.cfg0(max_steps=1)
    o0 = oracle_mod.OracleSim.from_spec(spec)
    params, state = o0.params, take_state(o0)
    for s in (0, 1):
        state[s]["h"] = np.full_like(state[s]["h"], 0.025)
    o, g = _one_step_both(oracle_mod, params, state, MIXED)
    for s in (0, 1):
        ro, rg = o.download(s), g.download(s)
        assert np.array_equal(rg["cell"], ro["cell"])
        assert np.array_equal(rg["nn"], ro["nn"])
        assert np.unique(ro["cell"], return_counts=True)[1].max() > 1024
        co, ho = o.neighbour_stats(s)
        cg, hg = g.neighbour_stats(s)
        assert np.array_equal(cg, co) and np.array_equal(hg, ho)
        act = ro["ind"] == 0
        e_rho = rel_pointwise(rg["rho"][act], ro["rho"][act])
        e_v = vec_normwise(rg, ro)
        print(f"\ndense species {s}: neighbours {float(cg.mean()):.0f} per particle, rho {e_rho:.3e} v {e_v:.3e}")
        assert e_rho <= 1e-5
        assert e_v <= 1e-5
        if s == 0:
            e_a = vec_normwise(rg, ro, ("ax", "ay", "az"))
            print(f"gas A {e_a:.3e}")
            assert e_a <= 1e-5


@pytest.mark.parametrize("spec_name", ["cfg0", "corner"])
def test_device_in_cell_order(oracle_mod, spec_name):
    """The device sorts by (cell, x at the start of the step, id): cells in the oracle's key
    order, and inside a cell the order of a lexicographic sort, so the rows of cells the pair
    kernels scan are x-sorted runs.  The reference's lists (kernel.cu:1937-1995) hold the same
    cells in index order, which the rebuilt nn lists match bit for bit."""
    spec = configs.cfg0() if spec_name == "cfg0" else configs.tube(0.04, 0.01, 1.0, 0.1, max_steps=6)
    params, state = _oracle_state(oracle_mod, spec, 3)
    print(f"\nin-cell order {spec_name}", flush=True)
    o, g = _one_step_both(oracle_mod, params, state, MIXED)
    for s in (0, 1):
        key = g.download(s)["cell"]
        ids = g.sorted_ids(s)
        n = len(key)
        assert np.array_equal(np.sort(ids), np.arange(n))
        expect = np.lexsort((np.arange(n), state[s]["x"], key))
        assert np.array_equal(ids, expect), f"species {s}: in-cell order differs"
        # the same cells in the same sequence as the oracle's (key, id) order
        assert np.array_equal(key[ids], key[o.sorted_ids(s)])
        # cells hold several particles here, so the x order is not the index order
        assert not np.array_equal(ids, o.sorted_ids(s))
