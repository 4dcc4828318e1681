"""The oracle's two-fluid scheme against the limits of the isothermal dusty shock tube (CPU).

With dust-to-gas ratio eps, gas and dust move together when the stopping time is far below the
flow time, as one isothermal fluid with sound speed c_s / sqrt(1 + eps); with negligible drag the
gas follows the plain isothermal Riemann solution (c_s) and the dust stays at rest.  A cfg0-sized
tube (spacing 0.005 / 0.01, eps = 1) run to t = 0.2 in FP64 checks both.  The bounds hold the
measured errors of this resolution with margin (coupled: star velocity -10.5%, L1 density 5.7%;
decoupled gas: star velocity -4.9%, L1 density 9%); swapping the two limits, dropping the drag
or the pressure kick fails them by far."""
import math

import numpy as np
import pytest

from dustyshock3d_amd import abi, configs, sim
from tests.parity_util import inject
from tests.riemann import shock_tube_errors

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


def run_tube(oracle_mod, K):
    spec = configs.tube(0.05, 0.005, K, 1.0, t_end=0.2, dt_mode=abi.DT_CFL, precision=abi.FP64)
    p = sim.tube_params(spec)
    o = inject(oracle_mod.OracleSim, p, {s: sim.tube_ic(spec, s) for s in (0, 1)})
    o.step(100000)
    st = o.status()
    assert st["done"] and abs(st["t"] - 0.2) < 1e-12
    return spec, st, o


def test_strongly_coupled_tube_moves_as_one_fluid(oracle_mod):
    spec, st, o = run_tube(oracle_mod, 1000.0)
    c = 1.0 / math.sqrt(2.0)
    for s in (0, 1):
        d = o.download(s)
        assert np.all(np.isfinite(d["vx"])) and np.all(np.isfinite(d["rho"]))
        err = shock_tube_errors(d, st["t"], spec.rho_left, c)
        print(f"\nspecies {s}: {err}")
        assert err["l1_rho"] <= 0.08 and err["l1_u"] <= 0.09, err
        assert abs(err["rho_star"]) <= 0.06 and abs(err["u_star"]) <= 0.13, err
        assert abs(err["shock"]) <= 0.03, err
        # and not the gas-alone solution
        alone = shock_tube_errors(d, st["t"], spec.rho_left, 1.0)
        assert alone["l1_u"] > 2 * err["l1_u"], (alone, err)


def test_uncoupled_gas_is_the_plain_riemann_solution(oracle_mod):
    spec, st, o = run_tube(oracle_mod, 0.001)
    gas, dust = o.download(0), o.download(1)
    err = shock_tube_errors(gas, st["t"], spec.rho_left, 1.0)
    print("\n", err)
    assert err["l1_rho"] <= 0.12 and err["l1_u"] <= 0.10, err
    assert abs(err["rho_star"]) <= 0.06 and abs(err["u_star"]) <= 0.08, err
    assert abs(err["shock"]) <= 0.025, err
    # the dust feels only the drag rate 1/t_stop = K / rho_d (<= 0.008 here): over t it picks
    # up at most t max|v_g| / min t_stop
    bound = st["t"] * np.max(np.abs(gas["vx"])) / np.min(dust["t_stop"])
    assert 0.0 < np.max(np.abs(dust["vx"])) <= bound, (np.max(np.abs(dust["vx"])), bound)
