"""Piece sizes of the tiled pair kernels and the autotuner that picks one: every piece size
gives the same results bit for bit (pieces split each candidate row in order), so tuning
never changes an answer."""
import numpy as np
import pytest

from dustyshock3d_amd import Simulation, abi, configs
from dustyshock3d_amd.batch import Batch

pytestmark = pytest.mark.gpu

TILES = (160, 128, 192, 224, 256)
FIELDS = ("x", "y", "z", "vx", "vy", "vz", "rho", "h", "e", "t_stop", "ax", "ay", "az", "edot", "nn", "cell")


@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_piece_sizes_agree_bit_for_bit(precision):
    spec = configs.cfg0(precision=precision, max_steps=20)
    runs = {}
    for tile in TILES:
        s = Simulation.from_spec(spec)
        s.set_tile(tile)
        s.step(20)
        runs[tile] = [s.download(sp) for sp in (0, 1)]
    for tile in TILES[1:]:
        for sp in (0, 1):
            for k in FIELDS:
                assert np.array_equal(runs[tile][sp][k], runs[TILES[0]][sp][k]), (tile, sp, k)


def test_autotune_keeps_the_answer():
    specs = [configs.tube(0.03, 0.005, 10.0, 1.0, max_steps=30, precision=abi.MIXED, tube_index=i) for i in range(3)]
    b = Batch(specs)
    res = b.autotune(steps_per_trial=2)
    print("\n", res)
    assert res["tile"] in TILES and res["variant"] == abi.PAIR_TILED and res["ms_per_step"] > 0
    b.run(100)
    ref = Batch(specs)
    ref.run(100)
    for i in range(len(specs)):
        assert b.status(i)["steps"] == ref.status(i)["steps"] == 30
        for sp in (0, 1):
            rb, rr = b.download(i, sp), ref.download(i, sp)
            for k in FIELDS:
                assert np.array_equal(rb[k], rr[k]), (i, sp, k)
