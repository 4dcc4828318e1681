"""Block shapes of the tiled pair kernels (piece size x waves per block) and the autotuner that
picks one: every piece size gives the same results bit for bit (pieces split each candidate
row in order), and so does every number of waves per block (each wave keeps its own chunk
and LDS piece), so tuning never changes an answer."""
import numpy as np
import pytest

from dustyshock3d_amd import Simulation, abi, configs
from dustyshock3d_amd.batch import Batch
from dustyshock3d_amd.sim import tube_ic, tube_params

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


# (piece size, waves per block); the first is the reference shape, the last the largest LDS
# footprint (the FP64 force kernel: four 256-candidate tiles, ~91 KB per block)
SHAPES = ((192, 1), (192, 2), (192, 4), (128, 4), (256, 2), (160, 4), (256, 4))


def _run_shape(spec, band, tile, waves, steps):
    p = tube_params(spec)
    p.fp64_wall_band = band
    s = Simulation(p)
    for sp in (0, 1):
        s.upload(sp, tube_ic(spec, sp))
    s.set_tile(tile)
    s.set_block_waves(waves)
    assert s.block_shape() == {"tile": tile, "waves": waves}
    s.set_stats(True)
    s.step(steps)
    stats = [s.neighbour_stats(sp) for sp in (0, 1)]
    s.set_stats(False)
    s.step(steps)
    return [s.download(sp) for sp in (0, 1)], stats


# FP64: the FP64 tiled kernels; mixed with the wall band: the fused kernels (FP64 band
# chunks first); mixed without it: the FP32 tiled kernels
@pytest.mark.parametrize("precision,band", [(abi.FP64, 1), (abi.MIXED, 1), (abi.MIXED, 0)])
def test_block_shapes_agree_bit_for_bit(precision, band):
    spec = configs.cfg0(precision=precision, max_steps=20)
    print(f"\nblock shapes, precision {precision} band {band}", flush=True)
    ref, ref_stats = _run_shape(spec, band, *SHAPES[0], 10)
    for tile, waves in SHAPES[1:]:
        got, stats = _run_shape(spec, band, tile, waves, 10)
        for sp in (0, 1):
            assert np.array_equal(stats[sp][0], ref_stats[sp][0]), (tile, waves, sp, "count")
            assert np.array_equal(stats[sp][1], ref_stats[sp][1]), (tile, waves, sp, "hash")
            for k in FIELDS:
                assert np.array_equal(got[sp][k], ref[sp][k]), (tile, waves, sp, k)


def test_autotune_keeps_the_answer():
    # 15 block shapes x (1 + 2) trial steps fit inside the 60 steps of each tube
    specs = [configs.tube(0.03, 0.005, 10.0, 1.0, max_steps=60, precision=abi.MIXED, tube_index=i) for i in range(3)]
    b = Batch(specs)
    res = b.autotune(steps_per_trial=2)
    print("\n", res)
    assert res["tile"] in TILES and res["waves"] in (1, 2, 4)
    assert res["variant"] == abi.PAIR_TILED and res["ms_per_step"] > 0
    assert b.block_shape() == {"tile": res["tile"], "waves": res["waves"]}
    b.run(100)
    ref = Batch(specs)
    ref.run(100)
    for i in range(len(specs)):
        assert b.status(i)["steps"] == ref.status(i)["steps"] == 60
        for sp in (0, 1):
            rb, rr = b.download(i, sp), ref.download(i, sp)
            for k in FIELDS:
                assert np.array_equal(rb[k], rr[k]), (i, sp, k)
