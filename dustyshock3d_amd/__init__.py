"""MI355X-native dusty-gas SPH step (gas + dust, implicit IDIC drag) behind a C ABI.

The compute path is libdsh.so (hand-written gfx950 HIP kernels + C++ host engine);
this package is the Python host mirror used by tests, the benchmark and scripts.
"""
from . import abi, configs  # noqa: F401
from .sim import (Simulation, default_params, read_init, tube_count, tube_ic,  # noqa: F401
                  tube_params)
