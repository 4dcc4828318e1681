"""Gas-only runs (n_dust = 0, the plain explicit scheme of kernel.cu:48-228) on the GPU
against the FP64 oracle and against the exact isothermal Riemann solution."""
import numpy as np
import pytest

from dustyshock3d_amd import Simulation
from dustyshock3d_amd.abi import FP64, MIXED, PAIR_DIRECT, PAIR_TILED
from tests.parity_util import inject, momentum_error, rel_normwise, rel_pointwise, take_state, vec_normwise
from tests.test_gas_only import check_riemann, gas_only_tube, riemann_errors

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [PAIR_TILED, PAIR_DIRECT])
@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_gas_only_step_parity(oracle_mod, precision, variant):
    spec, p, state = gas_only_tube()
    # a developed flow: 30 oracle steps, then one step of each from the same state
    o0 = inject(oracle_mod.OracleSim, p, state)
    o0.step(30)
    state = take_state(o0, (0,))
    p.precision = precision
    print(f"\ngas only precision={precision} variant={variant}", flush=True)
    o = inject(oracle_mod.OracleSim, p, state)
    g = inject(Simulation, p, state)
    g.set_pair_variant(variant)
    g.set_stats(True)
    o.step(1)
    g.step(1)
    ro, rg = o.download(0), g.download(0)
    assert np.array_equal(rg["cell"], ro["cell"]) and np.array_equal(rg["nn"], ro["nn"])
    co, ho = o.neighbour_stats(0)
    cg, hg = g.neighbour_stats(0)
    assert np.array_equal(cg, co) and np.array_equal(hg, ho)
    act = ro["ind"] == 0
    tol = 1e-10 if precision == FP64 else 1e-5
    e = {
        "rho": rel_pointwise(rg["rho"][act], ro["rho"][act]),
        "v": vec_normwise(rg, ro),
        "a": vec_normwise(rg, ro, ("ax", "ay", "az")),
        "edot": rel_normwise(rg["edot"], ro["edot"]),
        "e": rel_pointwise(rg["e"][act], ro["e"][act]),
    }
    print(e)
    assert max(e.values()) <= tol, e
    e_p = momentum_error(g.conservation(), o.conservation(), {0: ro})
    print(f"momentum {e_p:.3e}")
    assert e_p <= 1e-9


def test_gas_only_riemann_run():
    spec, p, state = gas_only_tube()
    g = inject(Simulation, p, state)
    g.step(100000)
    st = g.status()
    assert st["done"] and abs(st["t"] - 0.2) < 1e-12
    err = riemann_errors(g.download(0), st["t"], spec.rho_left)
    print("\n", err)
    check_riemann(err, 2.4 * spec.dx_left)
