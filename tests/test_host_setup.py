"""Host set-up of the library against the oracle's independent restatement (bit-exact),
and the Init.txt reader on the reference's own Init.txt (tests/golden/Init.txt)."""
import os

import numpy as np
import pytest

from dustyshock3d_amd import abi, configs, sim

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("spec", [configs.cfg0(), configs.sweep_tube(3), configs.sweep_tube(11),
                                  configs.tube(0.03, 0.01, 5.0, 0.3, tube_index=7)])
def test_ic_matches_oracle_bit_exact(oracle_mod, spec):
    assert sim.tube_count(spec) == oracle_mod.tube_count(spec)
    for s in (0, 1):
        a, b = sim.tube_ic(spec, s), oracle_mod.tube_ic(spec, s)
        for k in a:
            assert np.array_equal(a[k], b[k]), (s, k)
    pa = abi.params_dict(sim.tube_params(spec))
    pb = abi.params_dict(oracle_mod.tube_params(spec))
    assert pa == pb


def test_read_reference_init():
    p = sim.read_init(os.path.join(HERE, "golden", "Init.txt"))
    # Init.txt:2-18 in the order kernel.cu:1705-1722 reads them
    assert (p.x_min, p.x_max, p.y_min, p.y_max, p.z_min, p.z_max) == (-2.0, 2.0, -2.0, 2.0, -2.0, 2.0)
    assert p.eos_type == 0 and p.max_particles == 1000000
    assert p.particles_per_length == 40 and p.h == 0.1 and p.avg_cell_coeff == 0.5
    assert p.tau == 5.0e-3 and p.t_end == 0.2 and p.t_out == 0.1
    assert (p.alpha, p.beta, p.eps) == (1.0, 2.0, 0.01)
    # no extension lines: the reference constants
    assert p.gamma == 1.4 and p.cs == 1.0 and p.n_dust == 2 and p.h_mode == 0


def test_read_extended_init(tmp_path):
    base = open(os.path.join(HERE, "golden", "Init.txt")).read().rstrip("\n").split("\n")
    ext = ["1", "1", "1.4", "1.0", "1", "1.2", "1", "1", "100", "100", "1", "1", "0.5", "0.33", "1.0",
           "0", "2", "0.3", "0", "0", "0.05", "2", "1"]
    f = tmp_path / "Init.txt"
    f.write_text("\n".join(base + ext) + "\n")
    p = sim.read_init(str(f))
    assert p.periodic_y == 1 and p.h_mode == 1 and p.n_dust == 1 and p.K[0] == 100
    assert p.dt_mode == 2 and p.cfl == 0.3 and p.avg_factor == 2 and p.precision == 1


def test_sweep_distribution():
    specs = configs.cfg3(512)
    dx = np.array([s.dx_left for s in specs])
    K = np.array([s.K for s in specs])
    d2g = np.array([s.dust_to_gas for s in specs])
    assert dx.min() >= 0.0025 and dx.max() <= 0.01
    assert K.min() >= 0.01 and K.max() <= 1000
    assert d2g.min() >= 0.01 and d2g.max() <= 10
    n = np.array([sim.tube_count(s) for s in specs])
    assert n.min() < 3000 and n.max() > 60000
    # deterministic from the seed
    assert [s.K for s in configs.cfg3(4)] == list(K[:4])
