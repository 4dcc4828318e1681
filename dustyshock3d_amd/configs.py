"""The benchmark workloads as shock-tube specifications.

cfg0  plain FP64 tube, y,z in [0,0.05], dx 0.005/0.01, K=1, 50 fixed steps at CFL 0.3
cfg1  stiff tube, y,z in [0,0.1], dx 0.002/0.004, K=1000, mixed precision, 200 steps
cfg2  largest tube, y,z in [0,0.2], dx 0.002/0.004, K=100, mixed precision, to t=0.2
cfg3  sweep of 512 tubes: y,z in [0,0.05], left dx log-uniform in [0.0025,0.01],
      K log-uniform in [0.01,1000], dust-to-gas uniform in [0.01,10], to t=0.2
cfg4  serving stream: cfg3 tubes (95%) and cfg1-sized tubes (5%)
"""
from __future__ import annotations

import numpy as np

from .abi import DT_CFL, DT_CFL_ONCE, FP64, MIXED, TubeSpec

SEED = 0xD5D3


def tube(y_width, dx_left, K, dust_to_gas=1.0, *, t_end=1e30, dt_mode=DT_CFL_ONCE, max_steps=0,
         precision=FP64, tube_index=0, seed=SEED, cfl=0.3, jitter=1e-3, frozen_h=2.5,
         x_min=-0.5, x_max=0.5, rho_left=1.0) -> TubeSpec:
    return TubeSpec(x_min=x_min, x_max=x_max, y_width=y_width, z_width=y_width, dx_left=dx_left,
                    rho_left=rho_left, dust_to_gas=dust_to_gas, K=K, jitter=jitter,
                    frozen_h=frozen_h, t_end=t_end, cfl=cfl, dt_mode=dt_mode, max_steps=max_steps,
                    precision=precision, tube_index=tube_index, seed=seed)


def cfg0(**kw) -> TubeSpec:
    a = dict(max_steps=50, dt_mode=DT_CFL_ONCE, precision=FP64)
    a.update(kw)
    return tube(0.05, 0.005, 1.0, **a)


def cfg1(**kw) -> TubeSpec:
    a = dict(max_steps=200, dt_mode=DT_CFL_ONCE, precision=MIXED)
    a.update(kw)
    return tube(0.1, 0.002, 1000.0, **a)


def cfg2(**kw) -> TubeSpec:
    a = dict(t_end=0.2, dt_mode=DT_CFL, precision=MIXED)
    a.update(kw)
    return tube(0.2, 0.002, 100.0, **a)


def sweep_tube(i: int, seed: int = SEED, precision=MIXED, **kw) -> TubeSpec:
    """Tube i of the cfg3 sweep; every parameter comes from (seed, i)."""
    rng = np.random.default_rng([seed, i])
    dx = float(np.exp(rng.uniform(np.log(0.0025), np.log(0.01))))
    K = float(np.exp(rng.uniform(np.log(0.01), np.log(1000.0))))
    d2g = float(rng.uniform(0.01, 10.0))
    a = dict(t_end=0.2, dt_mode=DT_CFL, precision=precision, tube_index=i, seed=seed)
    a.update(kw)
    return tube(0.05, dx, K, d2g, **a)


def cfg3(n: int = 512, seed: int = SEED, **kw) -> list:
    return [sweep_tube(i, seed, **kw) for i in range(n)]


def large_tube(i: int, seed: int = SEED, precision=MIXED, **kw) -> TubeSpec:
    """The 5% large requests of cfg4 (cfg1-sized, 703,125 particles per species)."""
    rng = np.random.default_rng([seed, 7, i])
    K = float(np.exp(rng.uniform(np.log(0.01), np.log(1000.0))))
    d2g = float(rng.uniform(0.01, 10.0))
    a = dict(t_end=0.2, dt_mode=DT_CFL, precision=precision, tube_index=100000 + i, seed=seed)
    a.update(kw)
    return tube(0.1, 0.002, K, d2g, **a)
